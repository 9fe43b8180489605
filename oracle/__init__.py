"""FP64 CPU oracle for GooseFEM's element-quadrature + assembly path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import
this package, and only as the reference for comparisons.  The product (``goosefem_amd``) never imports it.

The numerics live in ``oracle.cpp`` (a loop-for-loop restatement of the reference files cited
there); this module wraps them with numpy.  Every function states the reference lines it follows.

Parity status of each part is recorded in ``oracle.cpp``'s header.  One part is
**parity unpinned** by the reference's own tests: the material law (``elastic_stress``,
``elastic_tangent``), which the reference delegates to the external GMat* packages; it is pinned
only by closed-form identities.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

QUAD4, QUAD4_PLANAR, HEX8 = 0, 1, 2
NNE = {QUAD4: 4, QUAD4_PLANAR: 4, HEX8: 8}
NDIM = {QUAD4: 2, QUAD4_PLANAR: 2, HEX8: 3}
TDIM = {QUAD4: 2, QUAD4_PLANAR: 3, HEX8: 3}


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


_i64 = ctypes.c_int64
_dbl = ctypes.c_double


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


# -------------------------------------------------------------------------------------------------
# Mesh: MeshQuad4.h:43-187, MeshHex8.h:41-609, Mesh.h:93-175,291-304,511-540,1577-1692,2462-2551
# -------------------------------------------------------------------------------------------------


def quad4_regular(nx: int, ny: int, h: float = 1.0):
    """(coor, conn) of Mesh::Quad4::Regular(nx, ny, h) (MeshQuad4.h:83-122)."""
    nnode = (nx + 1) * (ny + 1)
    coor = np.empty((nnode, 2))
    conn = np.empty((nx * ny, 4), dtype=np.int64)
    f = lib().or_quad4_regular
    f.argtypes = [_i64, _i64, _dbl, ctypes.c_void_p, ctypes.c_void_p]
    f(nx, ny, h, _p(coor), _p(conn))
    return coor, conn


def hex8_regular(nx: int, ny: int, nz: int, h: float = 1.0):
    """(coor, conn) of Mesh::Hex8::Regular(nx, ny, nz, h) (MeshHex8.h:77-131)."""
    nnode = (nx + 1) * (ny + 1) * (nz + 1)
    coor = np.empty((nnode, 3))
    conn = np.empty((nx * ny * nz, 8), dtype=np.int64)
    f = lib().or_hex8_regular
    f.argtypes = [_i64, _i64, _i64, _dbl, ctypes.c_void_p, ctypes.c_void_p]
    f(nx, ny, nz, h, _p(coor), _p(conn))
    return coor, conn


QUAD4_NODESETS = {
    "nodesBottomEdge": 0,
    "nodesTopEdge": 1,
    "nodesLeftEdge": 2,
    "nodesRightEdge": 3,
    "nodesBottomOpenEdge": 4,
    "nodesTopOpenEdge": 5,
    "nodesLeftOpenEdge": 6,
    "nodesRightOpenEdge": 7,
}


def quad4_nodeset(nx: int, ny: int, name: str) -> np.ndarray:
    """Edge node sets of Mesh::Quad4::Regular (MeshQuad4.h:128-167)."""
    f = lib().or_quad4_nodeset
    f.argtypes = [_i64, _i64, ctypes.c_int, ctypes.c_void_p]
    f.restype = _i64
    n = f(nx, ny, QUAD4_NODESETS[name], None)
    out = np.empty(n, dtype=np.int64)
    f(nx, ny, QUAD4_NODESETS[name], _p(out))
    return out


def quad4_nodes_periodic(nx: int, ny: int) -> np.ndarray:
    """nodesPeriodic of a 2-D regular mesh (Mesh.h:511-540)."""
    f = lib().or_quad4_nodes_periodic
    f.argtypes = [_i64, _i64, ctypes.c_void_p]
    f.restype = _i64
    n = f(nx, ny, None)
    out = np.empty((n, 2), dtype=np.int64)
    f(nx, ny, _p(out))
    return out


def hex8_nodes_periodic(nx: int, ny: int, nz: int) -> np.ndarray:
    """nodesPeriodic of a 3-D regular mesh (Mesh.h:1577-1692)."""
    f = lib().or_hex8_nodes_periodic
    f.argtypes = [_i64, _i64, _i64, ctypes.c_void_p]
    f.restype = _i64
    n = f(nx, ny, nz, None)
    out = np.empty((n, 2), dtype=np.int64)
    f(nx, ny, nz, _p(out))
    return out


def dofs(nnode: int, ndim: int) -> np.ndarray:
    """Mesh::dofs (Mesh.h:93-96)."""
    return np.arange(nnode * ndim, dtype=np.int64).reshape(nnode, ndim)


def dofs_periodic(nnode: int, ndim: int, ties: np.ndarray) -> np.ndarray:
    """RegularBase::dofsPeriodic (Mesh.h:291-304)."""
    ties = _i(ties)
    out = np.empty((nnode, ndim), dtype=np.int64)
    f = lib().or_dofs_periodic
    f.argtypes = [_i64, ctypes.c_int, ctypes.c_void_p, _i64, ctypes.c_void_p]
    f(nnode, ndim, _p(ties), ties.shape[0], _p(out))
    return out


def renumber(a: np.ndarray) -> np.ndarray:
    """Mesh::renumber (Mesh.h:114-175)."""
    a = _i(a)
    out = np.empty_like(a)
    f = lib().or_renumber
    f.argtypes = [ctypes.c_void_p, _i64, ctypes.c_void_p]
    f(_p(a), a.size, _p(out))
    return out


def coordination(conn: np.ndarray) -> np.ndarray:
    """Mesh::coordination (Mesh.h:2462-2477)."""
    conn = _i(conn)
    nv = int(conn.max()) + 1
    out = np.empty(nv, dtype=np.int64)
    f = lib().or_coordination
    f.argtypes = [ctypes.c_void_p, _i64, _i64, ctypes.c_void_p]
    f(_p(conn), conn.size, nv, _p(out))
    return out


def _csr_to_lists(ptr, idx):
    return [list(idx[ptr[i] : ptr[i + 1]]) for i in range(len(ptr) - 1)]


def elem2node(conn: np.ndarray, dofs_: np.ndarray | None = None):
    """Mesh::elem2node (Mesh.h:2485-2510), periodic-aware when dofs given (Mesh.h:2516-2551)."""
    conn = _i(conn)
    nelem, nne = conn.shape
    if dofs_ is None:
        nnode = int(conn.max()) + 1
        f = lib().or_elem2node
        f.argtypes = [ctypes.c_void_p, _i64, ctypes.c_int, _i64, ctypes.c_void_p, ctypes.c_void_p]
        f.restype = _i64
        nnz = f(_p(conn), nelem, nne, nnode, None, None)
        ptr = np.empty(nnode + 1, dtype=np.int64)
        idx = np.empty(nnz, dtype=np.int64)
        f(_p(conn), nelem, nne, nnode, _p(ptr), _p(idx))
    else:
        dofs_ = _i(dofs_)
        nnode, ndim = dofs_.shape
        f = lib().or_elem2node_periodic
        f.argtypes = [ctypes.c_void_p, _i64, ctypes.c_int, ctypes.c_void_p, _i64, ctypes.c_int,
                      ctypes.c_void_p, ctypes.c_void_p]
        f.restype = _i64
        nnz = f(_p(conn), nelem, nne, _p(dofs_), nnode, ndim, None, None)
        ptr = np.empty(nnode + 1, dtype=np.int64)
        idx = np.empty(nnz, dtype=np.int64)
        f(_p(conn), nelem, nne, _p(dofs_), nnode, ndim, _p(ptr), _p(idx))
    return _csr_to_lists(ptr, idx)


def nodaltyings(dofs_: np.ndarray):
    """Mesh::nodaltyings (Mesh.h:2839-2861)."""
    dofs_ = _i(dofs_)
    nnode, ndim = dofs_.shape
    f = lib().or_nodaltyings
    f.argtypes = [ctypes.c_void_p, _i64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = _i64
    nnz = f(_p(dofs_), nnode, ndim, None, None)
    ptr = np.empty(nnode + 1, dtype=np.int64)
    idx = np.empty(nnz, dtype=np.int64)
    f(_p(dofs_), nnode, ndim, _p(ptr), _p(idx))
    return _csr_to_lists(ptr, idx)


def node2dof(dofs_: np.ndarray):
    """Mesh::node2dof (Mesh.h:2560) = elem2node(dofs) (nodes per DOF, sorted)."""
    return elem2node(dofs_)


# -------------------------------------------------------------------------------------------------
# Quadrature (ElementQuad4.h, ElementQuad4Planar.h, ElementHex8.h, Element.h:892-1229)
# -------------------------------------------------------------------------------------------------


def quad_rule(et: int, nodal: bool = False):
    """Gauss (ElementQuad4.h:36-93, ElementHex8.h:26-110) or Nodal rule (xi, w)."""
    nd = NDIM[et]
    nip = 4 if nd == 2 else 8
    xi = np.empty((nip, nd))
    w = np.empty(nip)
    f = lib().or_quad_rule
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f(et, int(nodal), _p(xi), _p(w))
    return xi, w


class Quadrature:
    """FP64 restatement of Element::{Quad4::Quadrature, Quad4::QuadraturePlanar, Hex8::Quadrature}.

    ``et`` selects the element (QUAD4, QUAD4_PLANAR, HEX8).  ``x`` is the elemvec of nodal
    coordinates [nelem, nne, ndim].  ``faithful_q1`` reproduces the index quirk of the generic
    compute_dN (Element.h:1220) for Hex8.
    """

    def __init__(self, et, x, xi=None, w=None, thick=1.0, faithful_q1=False):
        self.et = et
        self.nne, self.ndim, self.tdim = NNE[et], NDIM[et], TDIM[et]
        if xi is None:
            xi, w = quad_rule(et)
        self.xi, self.w = _f64(xi), _f64(w)
        self.nip = self.w.size
        self.x = _f64(x)
        self.nelem = self.x.shape[0]
        self.thick = float(thick)
        self.N = np.empty((self.nip, self.nne))
        self.dNxi = np.empty((self.nip, self.nne, self.ndim))
        f = lib().or_quad_shape
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        f(et, self.nip, _p(self.xi), _p(self.N), _p(self.dNxi))
        self.faithful_q1 = faithful_q1
        self.update_x(self.x)

    def update_x(self, x):
        self.x = _f64(x)
        self.dNx = np.empty((self.nelem, self.nip, self.nne, self.ndim))
        self.vol = np.empty((self.nelem, self.nip))
        f = lib().or_quad_compute_dN
        f.argtypes = [ctypes.c_int, _i64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_void_p, _dbl, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        f(self.et, self.nelem, self.nip, _p(self.x), _p(self.dNxi), _p(self.w), self.thick,
          int(self.faithful_q1), _p(self.dNx), _p(self.vol))

    def GradN(self):
        return self.dNx

    def dV(self):
        return self.vol

    def _grad(self, mode, elemvec):
        elemvec = _f64(elemvec)
        out = np.empty((self.nelem, self.nip, self.tdim, self.tdim))
        f = lib().or_quad_grad
        f.argtypes = [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_void_p]
        f(self.et, mode, self.nelem, self.nip, _p(self.dNx), _p(elemvec), _p(out))
        return out

    def GradN_vector(self, elemvec):
        return self._grad(0, elemvec)

    def GradN_vector_T(self, elemvec):
        return self._grad(1, elemvec)

    def SymGradN_vector(self, elemvec):
        return self._grad(2, elemvec)

    def InterpQuad_vector(self, elemvec):
        elemvec = _f64(elemvec)
        n = elemvec.shape[2]
        out = np.empty((self.nelem, self.nip, n))
        f = lib().or_quad_interp_vector
        f.argtypes = [ctypes.c_int, _i64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_int, ctypes.c_void_p]
        f(self.et, self.nelem, self.nip, _p(self.N), _p(elemvec), n, _p(out))
        return out

    def Int_N_vector_dV(self, qvector):
        qvector = _f64(qvector)
        n = qvector.shape[2]
        out = np.empty((self.nelem, self.nne, n))
        f = lib().or_quad_int_N_vector_dV
        f.argtypes = [ctypes.c_int, _i64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        f(self.et, self.nelem, self.nip, _p(self.N), _p(self.vol), _p(qvector), n, _p(out))
        return out

    def Int_N_scalar_NT_dV(self, qscalar):
        qscalar = _f64(qscalar)
        K = self.nne * self.ndim
        out = np.empty((self.nelem, K, K))
        f = lib().or_quad_int_N_scalar_NT_dV
        f.argtypes = [ctypes.c_int, _i64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_void_p]
        f(self.et, self.nelem, self.nip, _p(self.N), _p(self.vol), _p(qscalar), _p(out))
        return out

    def Int_gradN_dot_tensor2_dV(self, qtensor):
        qtensor = _f64(qtensor)
        out = np.empty((self.nelem, self.nne, self.ndim))
        f = lib().or_quad_int_gradN_dot_tensor2_dV
        f.argtypes = [ctypes.c_int, _i64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_void_p]
        f(self.et, self.nelem, self.nip, _p(self.dNx), _p(self.vol), _p(qtensor), _p(out))
        return out

    def Int_gradN_dot_tensor4_dot_gradNT_dV(self, qtensor):
        qtensor = _f64(qtensor)
        K = self.nne * self.ndim
        out = np.empty((self.nelem, K, K))
        f = lib().or_quad_int_gradN_dot_tensor4_dot_gradNT_dV
        f.argtypes = [ctypes.c_int, _i64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_void_p]
        f(self.et, self.nelem, self.nip, _p(self.dNx), _p(self.vol), _p(qtensor), _p(out))
        return out


# -------------------------------------------------------------------------------------------------
# Vector (Vector.h:568-757), MatrixDiagonal (MatrixDiagonal.h:150-280), Matrix (Matrix.h:704-815)
# -------------------------------------------------------------------------------------------------


class Vector:
    """FP64 restatement of GooseFEM::Vector (serial loop order, last write wins on overwrite)."""

    def __init__(self, conn, dofs_):
        self.conn = _i(conn)
        self.dofs = _i(dofs_)
        self.nelem, self.nne = self.conn.shape
        self.nnode, self.ndim = self.dofs.shape
        self.ndof = int(self.dofs.max()) + 1

    def AsElement(self, arg):
        arg = _f64(arg)
        out = np.empty((self.nelem, self.nne, self.ndim))
        if arg.ndim == 2:
            _call("or_vec_as_element_nodevec", self.conn, self.nelem, _c(self.nne), _c(self.ndim),
                  arg, out)
        else:
            _call("or_vec_as_element_dofval", self.conn, self.dofs, self.nelem, _c(self.nne),
                  _c(self.ndim), arg, out)
        return out

    def AsDofs(self, arg):
        arg = _f64(arg)
        out = np.empty(self.ndof)
        if arg.ndim == 2:
            _call("or_vec_as_dofs_nodevec", self.dofs, self.nnode, _c(self.ndim), self.ndof, arg, out)
        else:
            _call("or_vec_as_dofs_elemvec", self.conn, self.dofs, self.nelem, _c(self.nne),
                  _c(self.ndim), self.ndof, arg, out)
        return out

    def AsNode(self, arg):
        arg = _f64(arg)
        out = np.empty((self.nnode, self.ndim))
        if arg.ndim == 1:
            _call("or_vec_as_node_dofval", self.dofs, self.nnode, _c(self.ndim), arg, out)
        else:
            _call("or_vec_as_node_elemvec", self.conn, self.nelem, _c(self.nne), self.nnode,
                  _c(self.ndim), arg, out)
        return out

    def AssembleDofs(self, arg):
        arg = _f64(arg)
        out = np.empty(self.ndof)
        if arg.ndim == 2:
            _call("or_vec_assemble_dofs_nodevec", self.dofs, self.nnode, _c(self.ndim), self.ndof,
                  arg, out)
        else:
            _call("or_vec_assemble_dofs_elemvec", self.conn, self.dofs, self.nelem, _c(self.nne),
                  _c(self.ndim), self.ndof, arg, out)
        return out

    def AssembleNode(self, arg):
        arg = _f64(arg)
        out = np.empty((self.nnode, self.ndim))
        _call("or_vec_assemble_node_elemvec", self.conn, self.dofs, self.nelem, _c(self.nne),
              self.nnode, _c(self.ndim), self.ndof, arg, out)
        return out


class _c(int):
    """Marks an argument passed as C int (not int64)."""


def _call(name, *args):
    """Call an oracle function: ndarray -> pointer, _c -> C int, int -> int64, float -> double."""
    fn = getattr(lib(), name)
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(_p(a))
        elif isinstance(a, _c):
            conv.append(ctypes.c_int(int(a)))
        elif isinstance(a, float):
            conv.append(_dbl(a))
        elif isinstance(a, (int, np.integer)):
            conv.append(_i64(int(a)))
        elif a is None:
            conv.append(None)
        else:
            conv.append(a)
    fn.restype = _i64
    return fn(*conv)


def diag_assemble(conn, dofs_, elemmat):
    """MatrixDiagonal::assemble (MatrixDiagonal.h:150-167): diagonal entries only, serial order."""
    conn, dofs_ = _i(conn), _i(dofs_)
    elemmat = _f64(elemmat)
    ndof = int(dofs_.max()) + 1
    out = np.empty(ndof)
    _call("or_diag_assemble", conn, dofs_, conn.shape[0], _c(conn.shape[1]), _c(dofs_.shape[1]),
          ndof, elemmat, out)
    return out


def diag_inverse(A):
    """MatrixDiagonal::factorize (MatrixDiagonal.h:268-280): inv = 1 / A."""
    A = _f64(A)
    out = np.empty_like(A)
    _call("or_diag_inverse", A.size, A, out)
    return out


def matrix_assemble(conn, dofs_, elemmat):
    """Matrix::assemble (Matrix.h:704-727) as CSR (rowptr, col, val), sorted columns."""
    conn, dofs_ = _i(conn), _i(dofs_)
    elemmat = None if elemmat is None else _f64(elemmat)
    ndof = int(dofs_.max()) + 1
    rowptr = np.empty(ndof + 1, dtype=np.int64)
    args = (conn, dofs_, conn.shape[0], _c(conn.shape[1]), _c(dofs_.shape[1]), ndof, elemmat)
    nnz = _call("or_matrix_assemble", *args, rowptr, None, None)
    col = np.empty(nnz, dtype=np.int64)
    val = np.empty(nnz)
    _call("or_matrix_assemble", *args, rowptr, col, val)
    return rowptr, col, val


def csr_dot(rowptr, col, val, x):
    """Matrix::Dot (Matrix.h:804-815): ascending-column row sums."""
    x = _f64(x)
    n = rowptr.size - 1
    b = np.empty(n)
    _call("or_csr_dot", n, _i(rowptr), _i(col), _f64(val), x, b)
    return b


# -------------------------------------------------------------------------------------------------
# Tyings (TyingsPeriodic.h:82-439)
# -------------------------------------------------------------------------------------------------


def tyings_control(coor, dofs_):
    """Tyings::Control (TyingsPeriodic.h:361-392): (coor, dofs, controlDofs, controlNodes)."""
    coor, dofs_ = _f64(coor), _i(dofs_)
    nnode, ndim = coor.shape
    co = np.empty((nnode + ndim, ndim))
    do = np.empty((nnode + ndim, ndim), dtype=np.int64)
    cd = np.empty((ndim, ndim), dtype=np.int64)
    cn = np.empty(ndim, dtype=np.int64)
    _call("or_tyings_control", coor, dofs_, nnode, _c(ndim), co, do, cd, cn)
    return co, do, cd, cn


def tyings_periodic(coor, dofs_, control_dofs, ties, iip):
    """Tyings::Periodic (TyingsPeriodic.h:82-121) + Cdi/Cdu/Cdp (:238-343).

    Returns dict with reordered dofs, control, nnu/nnp/nni/nnd and the three tying matrices as
    dense numpy arrays (small meshes) plus CSR triples.
    """
    coor, dofs_ = _f64(coor), _i(dofs_)
    control_dofs, ties, iip = _i(control_dofs), _i(ties), _i(iip)
    nnode, ndim = dofs_.shape
    dout = np.empty_like(dofs_)
    cout = np.empty_like(control_dofs)
    sizes = np.empty(4, dtype=np.int64)
    _call("or_tyings_periodic", dofs_, nnode, _c(ndim), control_dofs, ties, ties.shape[0], iip,
          iip.size, dout, cout, sizes)
    nnu, nnp, nni, nnd = (int(s) for s in sizes)
    res = dict(dofs=dout, control=cout, nnu=nnu, nnp=nnp, nni=nni, nnd=nnd)
    for which, name, ncol in ((0, "Cdi", nni), (1, "Cdu", nnu), (2, "Cdp", nnp)):
        rowptr = np.empty(nnd + 1, dtype=np.int64)
        args = (_c(which), coor, dout, _c(ndim), cout, ties, ties.shape[0], nnu)
        nnz = _call("or_tyings_C", *args, rowptr, None, None)
        col = np.empty(nnz, dtype=np.int64)
        val = np.empty(nnz)
        _call("or_tyings_C", *args, rowptr, col, val)
        res[name + "_csr"] = (rowptr, col, val)
        dense = np.zeros((nnd, ncol))
        for r in range(nnd):
            for k in range(rowptr[r], rowptr[r + 1]):
                dense[r, col[k]] += val[k]
        res[name] = dense
    return res


# -------------------------------------------------------------------------------------------------
# Material law (parity unpinned: external GMat*), velocity-Verlet, CG, input generator
# -------------------------------------------------------------------------------------------------


def elastic_stress(eps, K, G):
    """sigma = K tr(eps) I + 2G dev(eps) per point; K, G broadcast to eps.shape[:-2]."""
    eps = _f64(eps)
    td = eps.shape[-1]
    nq = eps.size // (td * td)
    Kq = _f64(np.broadcast_to(K, eps.shape[:-2]))
    Gq = _f64(np.broadcast_to(G, eps.shape[:-2]))
    out = np.empty_like(eps)
    _call("or_elastic_stress", nq, _c(td), eps, Kq, Gq, out)
    return out


def elastic_tangent(K, G, td=3):
    """C_ijkl = K d_ij d_kl + 2G (I_sym - I/3 d_ij d_kl) per point."""
    K = _f64(K)
    G = _f64(np.broadcast_to(G, K.shape))
    out = np.empty(K.shape + (td,) * 4)
    _call("or_elastic_tangent", K.size, _c(td), K, G, out)
    return out


def plastic_refresh(Eps, K, G, sigy0, H, Epsp_t, epsp_t, tangent=True):
    """J2 radial return with linear isotropic hardening (or_plastic_refresh; the law of
    GMatElastoPlastic.Cartesian3d.LinearHardening2d used at periodic_elastoplastic.py:71-104).
    Eps, Epsp_t: [..., 3, 3]; K, G, sigy0, H, epsp_t: [...].  Returns (Sig, C or None, Epsp, epsp)."""
    Eps = _f64(Eps)
    shape = Eps.shape[:-2]
    nq = int(np.prod(shape)) if shape else 1
    args = [_f64(np.broadcast_to(a, shape)) for a in (K, G, sigy0, H)]
    Epsp_t, epsp_t = _f64(Epsp_t), _f64(epsp_t)
    Sig = np.empty_like(Eps)
    Epsp = np.empty_like(Eps)
    epsp = np.empty(shape)
    C = np.empty(shape + (3, 3, 3, 3)) if tangent else None
    _call("or_plastic_refresh", nq, Eps, *args, Epsp_t, epsp_t, Epsp, epsp, Sig, C)
    return Sig, C, Epsp, epsp


class LinearHardening:
    """The state machine of GMatElastoPlastic.Cartesian3d.LinearHardening2d as the statics example
    drives it (periodic_elastoplastic.py:85-104): set Eps, refresh(), read Sig / C; increment()
    commits the plastic state."""

    def __init__(self, K, G, sigy0, H):
        self.K, self.G, self.sigy0, self.H = (_f64(a) for a in (K, G, sigy0, H))
        shape = self.K.shape
        self.Eps = np.zeros(shape + (3, 3))
        self.Epsp = np.zeros(shape + (3, 3))
        self.Epsp_t = np.zeros(shape + (3, 3))
        self.epsp = np.zeros(shape)
        self.epsp_t = np.zeros(shape)
        self.Sig = np.zeros(shape + (3, 3))
        self.C = None

    def refresh(self, tangent=True):
        self.Sig, self.C, self.Epsp, self.epsp = plastic_refresh(
            self.Eps, self.K, self.G, self.sigy0, self.H, self.Epsp_t, self.epsp_t, tangent)

    def increment(self):
        self.Epsp_t = self.Epsp.copy()
        self.epsp_t = self.epsp.copy()


def vv_steps(nsteps, et, conn, dofs_, dNx, vol, Ke, Ge, inv_mass, dt, u, v, a, fext=None):
    """Velocity-Verlet steps (Elastic-VelocityVerlet/main.cpp:110-157); u, v, a updated in place.

    Returns fint (dofval) of the last step.
    """
    conn, dofs_ = _i(conn), _i(dofs_)
    ndof = int(dofs_.max()) + 1
    fint = np.empty(ndof)
    c = 0.5 * dt * dt
    for arr in (u, v, a):
        assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    _call("or_vv_steps", int(nsteps), _c(et), conn, dofs_, conn.shape[0], ndof, _f64(dNx), _f64(vol),
          _f64(Ke), _f64(Ge), _f64(inv_mass), None if fext is None else _f64(fext), float(dt),
          float(c), u, v, a, fint)
    return fint


def vv_steps_parallel(nsteps, et, conn, dofs_, dNx, vol, Ke, Ge, inv_mass, dt, u, v, a, fext=None):
    """vv_steps on all host threads (OpenMP; the all-cores CPU baseline), bit-identical to vv_steps.

    Returns (fint of the last step, threads used).
    """
    conn, dofs_ = _i(conn), _i(dofs_)
    ndof = int(dofs_.max()) + 1
    fint = np.empty(ndof)
    c = 0.5 * dt * dt
    for arr in (u, v, a):
        assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    nt = _call("or_vv_steps_parallel", int(nsteps), _c(et), conn, dofs_, conn.shape[0], ndof, _f64(dNx),
               _f64(vol), _f64(Ke), _f64(Ge), _f64(inv_mass), None if fext is None else _f64(fext), float(dt),
               float(c), u, v, a, fint)
    return fint, int(nt)


def pcg_csr(rowptr, col, val, diag, b, x0=None, tol=1e-10, maxiter=100000):
    """Plain FP64 Jacobi-PCG (not in the reference). Returns (x, iterations, true rel. residual)."""
    n = rowptr.size - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    res = np.zeros(1)
    it = _call("or_pcg_csr", n, _i(rowptr), _i(col), _f64(val), _f64(diag), _f64(b), x, float(tol),
               int(maxiter), res)
    return x, int(it), float(res[0])


def randn(seed: int, n: int, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
    """SplitMix64 -> Box-Muller portable normal generator."""
    out = np.empty(n)
    f = lib().or_randn
    f.argtypes = [ctypes.c_uint64, _i64, _dbl, _dbl, ctypes.c_void_p]
    f(seed, n, mean, std, _p(out))
    return out


# -------------------------------------------------------------------------------------------------
# Dirichlet partitioning: VectorPartitioned.h:72-661, MatrixDiagonalPartitioned.h:42-452.  Plain
# loops in the reference's loop order (small meshes only).
# -------------------------------------------------------------------------------------------------


class VectorPartitioned:
    """GooseFEM::VectorPartitioned: iiu = setdiff1d(dofs, iip) (VectorPartitioned.h:80),
    part = Mesh::Reorder({iiu, iip}).apply(dofs) (Mesh.h:2384-2454)."""

    def __init__(self, conn, dofs_, iip):
        self.conn, self.dofs = _i(conn), _i(dofs_)
        self.iip = _i(iip).ravel()
        assert np.unique(self.iip).size == self.iip.size
        self.nelem, self.nne = self.conn.shape
        self.nnode, self.ndim = self.dofs.shape
        self.ndof = int(self.dofs.max()) + 1
        self.iiu = np.setdiff1d(self.dofs, self.iip)  # sorted unique, like xt::setdiff1d
        self.nnu, self.nnp = self.iiu.size, self.iip.size
        renum = np.full(max(self.ndof, int(self.iip.max()) + 1 if self.nnp else 0), -1, dtype=np.int64)
        i = 0
        for arr in (self.iiu, self.iip):  # Reorder: running index over the concatenation
            for j in arr:
                renum[j] = i
                i += 1
        self.part = renum[self.dofs]

    def dofs_is_u(self):  # VectorPartitioned.h:126-143
        return self.part < self.nnu

    def dofs_is_p(self):  # VectorPartitioned.h:147-161
        return self.part >= self.nnu

    def _as_dofs(self, arg, want_u):
        """asDofs_u / asDofs_p (VectorPartitioned.h:443-661): zero fill, then overwrite in loop order."""
        arg = _f64(arg)
        n = self.nnu if want_u else self.nnp
        out = np.zeros(n)
        sel = (lambda d: d < self.nnu) if want_u else (lambda d: d >= self.nnu)
        off = 0 if want_u else self.nnu
        if arg.ndim == 1:
            idx = self.iiu if want_u else self.iip
            for d in range(n):
                out[d] = arg[idx[d]]
        elif arg.ndim == 2:
            for m in range(self.nnode):
                for i in range(self.ndim):
                    d = self.part[m, i]
                    if sel(d):
                        out[d - off] = arg[m, i]
        else:
            for e in range(self.nelem):
                for m in range(self.nne):
                    for i in range(self.ndim):
                        d = self.part[self.conn[e, m], i]
                        if sel(d):
                            out[d - off] = arg[e, m, i]
        return out

    def AsDofs_u(self, arg):
        return self._as_dofs(arg, True)

    def AsDofs_p(self, arg):
        return self._as_dofs(arg, False)

    def DofsFromParitioned(self, u, p):  # VectorPartitioned.h:293-307
        out = np.zeros(self.ndof)
        for d in range(self.nnu):
            out[self.iiu[d]] = u[d]
        for d in range(self.nnp):
            out[self.iip[d]] = p[d]
        return out

    def _pick(self, d, u, p):
        return u[d] if d < self.nnu else p[d - self.nnu]

    def NodeFromPartitioned(self, u, p):  # VectorPartitioned.h:344-358
        out = np.empty((self.nnode, self.ndim))
        for m in range(self.nnode):
            for i in range(self.ndim):
                out[m, i] = self._pick(self.part[m, i], u, p)
        return out

    def ElementFromPartitioned(self, u, p):  # VectorPartitioned.h:395-412
        out = np.empty((self.nelem, self.nne, self.ndim))
        for e in range(self.nelem):
            for m in range(self.nne):
                for i in range(self.ndim):
                    out[e, m, i] = self._pick(self.part[self.conn[e, m], i], u, p)
        return out

    def Copy_u(self, src, dest):  # VectorPartitioned.h:198-212
        out = _f64(dest).copy()
        mask = self.dofs_is_u()
        out[mask] = _f64(src)[mask]
        return out

    def Copy_p(self, src, dest):  # VectorPartitioned.h:251-263
        out = _f64(dest).copy()
        mask = self.dofs_is_p()
        out[mask] = _f64(src)[mask]
        return out


class MatrixDiagonalPartitioned:
    """GooseFEM::MatrixDiagonalPartitioned (MatrixDiagonalPartitioned.h:42-452)."""

    def __init__(self, conn, dofs_, iip):
        self.v = VectorPartitioned(conn, dofs_, iip)
        self.Auu = np.zeros(self.v.nnu)
        self.App = np.zeros(self.v.nnp)

    def assemble(self, elemmat):  # MatrixDiagonalPartitioned.h:73-98
        v = self.v
        elemmat = _f64(elemmat)
        self.Auu[:] = 0.0
        self.App[:] = 0.0
        for e in range(v.nelem):
            for m in range(v.nne):
                for i in range(v.ndim):
                    d = v.part[v.conn[e, m], i]
                    r = m * v.ndim + i
                    if d < v.nnu:
                        self.Auu[d] += elemmat[e, r, r]
                    else:
                        self.App[d - v.nnu] += elemmat[e, r, r]

    def set(self, A):  # MatrixDiagonalPartitioned.h:121-135
        A = _f64(A)
        self.Auu = A[self.v.iiu].copy()
        self.App = A[self.v.iip].copy()

    def data(self):  # MatrixDiagonalPartitioned.h:142-156
        out = np.zeros(self.v.ndof)
        out[self.v.iiu] = self.Auu
        out[self.v.iip] = self.App
        return out

    def _diag_of(self, d):
        return self.Auu[d] if d < self.v.nnu else self.App[d - self.v.nnu]

    def Dot(self, x):  # MatrixDiagonalPartitioned.h:189-224
        v = self.v
        x = _f64(x)
        b = np.zeros_like(x)  # DOF numbers absent from dofs (tied-away holes) are left 0 here
        if x.ndim == 2:
            for m in range(v.nnode):
                for i in range(v.ndim):
                    b[m, i] = self._diag_of(v.part[m, i]) * x[m, i]
        else:
            for d in range(v.nnu):
                b[v.iiu[d]] = self.Auu[d] * x[v.iiu[d]]
            for d in range(v.nnp):
                b[v.iip[d]] = self.App[d] * x[v.iip[d]]
        return b

    def Dot_u(self, x_u, x_p):  # :248-264
        return self.Auu * _f64(x_u)

    def Dot_p(self, x_u, x_p):  # :286-300
        return self.App * _f64(x_p)

    def Solve(self, b, x):  # :306-337 (x: the values kept in the prescribed entries)
        v = self.v
        inv_uu = 1.0 / self.Auu  # factorize, :440-452
        b, out = _f64(b), _f64(x).copy()
        if b.ndim == 2:
            for m in range(v.nnode):
                for i in range(v.ndim):
                    if v.part[m, i] < v.nnu:
                        out[m, i] = inv_uu[v.part[m, i]] * b[m, i]
        else:
            for d in range(v.nnu):
                out[v.iiu[d]] = inv_uu[d] * b[v.iiu[d]]
        return out

    def Solve_u(self, b_u, x_p):  # :356-372
        return (1.0 / self.Auu) * _f64(b_u)

    def Reaction(self, x, b):  # :378-403 via Matrix.h:490-514
        v = self.v
        x, out = _f64(x), _f64(b).copy()
        if x.ndim == 2:
            for m in range(v.nnode):
                for i in range(v.ndim):
                    if v.part[m, i] >= v.nnu:
                        out[m, i] = self.App[v.part[m, i] - v.nnu] * x[m, i]
        else:
            for d in range(v.nnp):
                out[v.iip[d]] = self.App[d] * x[v.iip[d]]
        return out

    def Reaction_p(self, x_u, x_p):  # :405-420
        return self.App * _f64(x_p)


class MatrixPartitioned:
    """GooseFEM::MatrixPartitioned (MatrixPartitioned.h:34-523) and the solve of
    MatrixPartitionedSolver (:526-640), restated with pure-Python loops (small meshes only).

    Blocks 0..3 = uu, up, pu, pp, each as {(row, col): value} in the partitioned numbering."""

    def __init__(self, conn, dofs_, iip):
        self.v = VectorPartitioned(conn, dofs_, iip)
        self.blocks = [{}, {}, {}, {}]

    def _block_of(self, di, dj):
        nu = self.v.nnu
        b = (2 if di >= nu else 0) + (1 if dj >= nu else 0)
        return b, (di - nu if di >= nu else di, dj - nu if dj >= nu else dj)

    def assemble(self, elemmat):
        """:156-210: triplets in (e, m, i, n, j) order; setFromTriplets sums duplicates in that order."""
        v = self.v
        elemmat = _f64(elemmat)
        self.blocks = [{}, {}, {}, {}]
        nd = v.ndim
        for e in range(v.nelem):
            for m in range(v.nne):
                for i in range(nd):
                    di = v.part[v.conn[e, m], i]
                    for n in range(v.nne):
                        for j in range(nd):
                            dj = v.part[v.conn[e, n], j]
                            b, key = self._block_of(di, dj)
                            val = elemmat[e, m * nd + i, n * nd + j]
                            blk = self.blocks[b]
                            blk[key] = blk[key] + val if key in blk else val

    def csr(self, b):
        """Block b as CSR (rowptr, col, val): rows and columns ascending."""
        nrows = self.v.nnu if b < 2 else self.v.nnp
        keys = sorted(self.blocks[b])
        rowptr = np.zeros(nrows + 1, dtype=np.int64)
        for r, _ in keys:
            rowptr[r + 1] += 1
        rowptr = np.cumsum(rowptr)
        col = np.array([c for _, c in keys], dtype=np.int64)
        val = np.array([self.blocks[b][k] for k in keys], dtype=np.float64)
        return rowptr, col, val

    def _rows(self, b):
        nrows = self.v.nnu if b < 2 else self.v.nnp
        rows = [[] for _ in range(nrows)]
        for (r, c) in sorted(self.blocks[b]):
            rows[r].append((c, self.blocks[b][(r, c)]))
        return rows

    def _product(self, first, second, xa, xb, init=None, sign=1.0):
        """Row sums over ascending columns of block `first`, then of block `second`, one running sum
        per row: Eigen's "dst = A xa; dst += B xb" for column-major A, B (SparseDenseProduct's
        res(i) += value * rhs(j), columns in order)."""
        ra, rb = self._rows(first), self._rows(second)
        out = np.empty(len(ra))
        for r in range(len(ra)):
            s = 0.0 if init is None else init[r]
            for c, a in ra[r]:
                s += a * (sign * xa[c])
            for c, a in rb[r]:
                s += a * (sign * xb[c])
            out[r] = s
        return out

    def Dot_u(self, x_u, x_p):  # B_u of :378-396
        return self._product(0, 1, _f64(x_u), _f64(x_p))

    def Dot_p(self, x_u, x_p):  # B_p of :378-396; Reaction_p, :436-447
        return self._product(2, 3, _f64(x_u), _f64(x_p))

    Reaction_p = Dot_p

    def Todense(self):  # :324-350
        v = self.v
        out = np.zeros((v.ndof, v.ndof))
        rows = (v.iiu, v.iiu, v.iip, v.iip)
        cols = (v.iiu, v.iip, v.iiu, v.iip)
        for b in range(4):
            for (r, c), a in self.blocks[b].items():
                out[rows[b][r], cols[b][c]] = a
        return out

    def Dot(self, x):  # :354-396 (entries of DOF numbers outside iiu and iip: 0)
        v = self.v
        x = _f64(x)
        bu = self.Dot_u(v.AsDofs_u(x), v.AsDofs_p(x))
        bp = self.Dot_p(v.AsDofs_u(x), v.AsDofs_p(x))
        if x.ndim == 2:
            return v.NodeFromPartitioned(bu, bp)
        return v.DofsFromParitioned(bu, bp)

    def Reaction(self, x, b):  # :401-433 via Matrix.h:490-514
        v = self.v
        x, out = _f64(x), _f64(b).copy()
        bp = self.Dot_p(v.AsDofs_u(x), v.AsDofs_p(x))
        if x.ndim == 2:
            for m in range(v.nnode):
                for i in range(v.ndim):
                    if v.part[m, i] >= v.nnu:
                        out[m, i] = bp[v.part[m, i] - v.nnu]
        else:
            out[v.iip] = bp
        return out

    def Solve_u(self, b_u, x_p):
        """:615-631: x_u = A_uu^-1 (b_u - A_up x_p); the dense LU solve stands in for the sparse LDLT
        (both are direct solves: they agree to rounding)."""
        v = self.v
        rows = self._rows(1)
        rhs = np.empty(v.nnu)
        b_u, x_p = _f64(b_u), _f64(x_p)
        for r in range(v.nnu):
            s = b_u[r]
            for c, a in rows[r]:
                s -= a * x_p[c]
            rhs[r] = s
        Auu = np.zeros((v.nnu, v.nnu))
        for (r, c), a in self.blocks[0].items():
            Auu[r, c] = a
        return np.linalg.solve(Auu, rhs) if v.nnu else rhs

    def Solve(self, b, x):  # :583-613: unknown entries of a copy of x
        v = self.v
        b, out = _f64(b), _f64(x).copy()
        xu = self.Solve_u(v.AsDofs_u(b), v.AsDofs_p(out))
        if b.ndim == 2:
            for m in range(v.nnode):
                for i in range(v.ndim):
                    if v.part[m, i] < v.nnu:
                        out[m, i] = xu[v.part[m, i]]
        else:
            out[v.iiu] = xu
        return out


class QuadratureAxisymmetric:
    """Element::Quad4::QuadratureAxisymmetric (ElementQuad4Axisymmetric.h:33-467): nodal coordinates
    (z, r), tensors in (r, theta, z) order; G[e, q, m] = (dN/dr, N / r_q, dN/dz)."""

    def __init__(self, x, xi=None, w=None):
        if xi is None:
            xi, w = quad_rule(QUAD4)
        self.xi, self.w = _f64(xi), _f64(w)
        self.nip, self.nne, self.ndim, self.tdim = self.w.size, 4, 2, 3
        self.N = np.empty((self.nip, 4))
        self.dNxi = np.empty((self.nip, 4, 2))
        f = lib().or_quad_shape
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        f(QUAD4, self.nip, _p(self.xi), _p(self.N), _p(self.dNxi))
        self.update_x(x)

    def update_x(self, x):
        self.x = _f64(x)
        self.nelem = self.x.shape[0]
        self.G = np.empty((self.nelem, self.nip, 4, 3))
        self.vol = np.empty((self.nelem, self.nip))
        _call("or_axi_geometry", self.nelem, _c(self.nip), self.x, self.N, self.dNxi, self.w, self.G, self.vol)

    def dV(self):
        return self.vol

    def B(self):
        """The reference's B [nelem, nip, nne, 3, 3, 3] (ElementQuad4Axisymmetric.h:401-438)."""
        B = np.zeros((self.nelem, self.nip, 4, 3, 3, 3))
        B[:, :, :, 0, 0, 0] = self.G[..., 0]
        B[:, :, :, 0, 2, 2] = self.G[..., 0]
        B[:, :, :, 1, 1, 0] = self.G[..., 1]
        B[:, :, :, 2, 0, 0] = self.G[..., 2]
        B[:, :, :, 2, 2, 2] = self.G[..., 2]
        return B

    def _grad(self, mode, elemvec):
        out = np.empty((self.nelem, self.nip, 3, 3))
        _call("or_axi_grad", _c(mode), self.nelem, _c(self.nip), self.G, _f64(elemvec), out)
        return out

    def GradN_vector(self, elemvec):
        return self._grad(0, elemvec)

    def GradN_vector_T(self, elemvec):
        return self._grad(1, elemvec)

    def SymGradN_vector(self, elemvec):
        return self._grad(2, elemvec)

    def Int_N_scalar_NT_dV(self, qscalar):
        out = np.empty((self.nelem, 8, 8))
        _call("or_axi_int_N_scalar_NT_dV", self.nelem, _c(self.nip), self.N, self.vol, _f64(qscalar), out)
        return out

    def Int_gradN_dot_tensor2_dV(self, qtensor):
        out = np.empty((self.nelem, 4, 2))
        _call("or_axi_int_tensor2", self.nelem, _c(self.nip), self.G, self.vol, _f64(qtensor), out)
        return out

    def Int_gradN_dot_tensor4_dot_gradNT_dV(self, qtensor):
        out = np.empty((self.nelem, 8, 8))
        _call("or_axi_int_tensor4", self.nelem, _c(self.nip), self.G, self.vol, _f64(qtensor), out)
        return out
