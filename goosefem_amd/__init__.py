"""goosefem_amd: MI355X-native element-quadrature and assembly path of GooseFEM.

Python mirror of the reference's Python API (python/*.hpp bindings of include/GooseFEM) over the
C ABI declared in include/goosefem_amd.h.  Device arrays are torch tensors on the GPU (torch is
used for device memory and streams only); every computation runs in libgoosefem_amd.so.  There is
no CPU fallback: if the library is missing, importing the compute classes raises.

    import goosefem_amd as gf
    mesh = gf.Mesh.Quad4.Regular(64, 64)
    vec = gf.Vector(mesh.conn, mesh.dofs)
    quad = gf.Element.Quad4.QuadraturePlanar(vec.AsElement(gf.to_device(mesh.coor)))
    eps = quad.SymGradN_vector(vec.AsElement(u))
"""

from __future__ import annotations

import ctypes
import os
from types import SimpleNamespace

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgoosefem_amd.so")

GF_OK, GF_EINVAL, GF_EHIP, GF_ENOMEM, GF_ESTATE, GF_ENOTCONV = 0, -1, -2, -3, -4, -5
QUAD4, QUAD4_PLANAR, HEX8 = 0, 1, 2
GAUSS, NODAL = 0, 1
F64, F32, BF16X3, F64_FAST = 0, 1, 2, 3
GF_START_GIVEN, GF_START_PREVIOUS, GF_START_ZERO = 0, 1, 2

_lib = None


class GooseFEMError(RuntimeError):
    """Raised for a nonzero status of the C ABI (the reference raises std::runtime_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _hip_runtimes_mapped():
    try:
        maps = open("/proc/self/maps").read().splitlines()
    except OSError:
        return set()
    return {ln.split()[-1] for ln in maps if "libamdhip64" in ln}


def lib() -> ctypes.CDLL:
    """The loaded C-ABI library.  Raises if it has not been built.

    torch is imported first so that the library's libamdhip64.so.7 dependency binds (by SONAME)
    to the HIP runtime torch already loaded: one runtime per process, so device pointers and
    streams are shared.  Two mapped runtimes raise instead of failing later.
    """
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build it with `python -m goosefem_amd.build` "
                "(there is no CPU fallback)"
            )
        import torch  # noqa: F401  (binds the HIP runtime before the library loads)

        _lib = ctypes.CDLL(LIB_PATH)
        _lib.gf_last_error.restype = ctypes.c_char_p
        rts = _hip_runtimes_mapped()
        if len(rts) > 1:
            raise ImportError(f"two HIP runtimes mapped in one process: {sorted(rts)}")
    return _lib


def _check(rc: int):
    if rc != GF_OK:
        raise GooseFEMError(rc, lib().gf_last_error().decode())


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f64 = ctypes.c_double


_CTYPE_OF = {
    "int": ctypes.c_int,
    "int32_t": ctypes.c_int32,
    "int64_t": ctypes.c_int64,
    "uint64_t": ctypes.c_uint64,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
}
_signatures = None


def signatures():
    """{name: [ctypes argument types]} of every function the C ABI header declares.

    Parsed from include/goosefem_amd.h, so that each call is checked for its argument count and
    scalar types before it crosses the ABI (a mismatch raises TypeError instead of passing garbage
    through a register).  Pointers and handles are void*, `const char*` is a C string.
    """
    global _signatures
    if _signatures is None:
        import re

        hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "goosefem_amd.h")
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        sigs = {}
        for m in re.finditer(r"\bint\s+(gf_\w+)\s*\(([^)]*)\)\s*;", text):
            name, params = m.group(1), " ".join(m.group(2).split())
            types = []
            if params not in ("", "void"):
                for p in params.split(","):
                    p = p.strip()
                    if "*" in p:
                        is_str = re.match(r"const\s+char\s*\*\s*\w+$", p) is not None
                        types.append(ctypes.c_char_p if is_str else ctypes.c_void_p)
                        continue
                    words = p.replace("const ", "").split()
                    base = words[0]
                    if base in _CTYPE_OF:
                        types.append(_CTYPE_OF[base])
                    elif base.startswith("gf_"):  # opaque handle
                        types.append(ctypes.c_void_p)
                    else:
                        raise ImportError(f"goosefem_amd.h: unknown parameter type {p!r} of {name}")
            sigs[name] = types
        _signatures = sigs
    return _signatures


def _fn(name):
    """The ABI function `name` with its argument types and int status return set."""
    f = getattr(lib(), name)
    if getattr(f, "_gf_typed", False) is False:
        sig = signatures().get(name)
        if sig is None:
            raise AttributeError(f"{name} is not declared in include/goosefem_amd.h")
        f.argtypes = sig
        f.restype = ctypes.c_int
        f._gf_typed = True
    return f


def _call(name, *args):
    f = _fn(name)
    if len(args) != len(f.argtypes):
        raise TypeError(f"{name} takes {len(f.argtypes)} arguments ({len(args)} given)")
    _check(f(*args))


def _hp(a: np.ndarray):
    return a.ctypes.data_as(_vp)


# -------------------------------------------------------------------------------------------------
# device memory helpers (torch)
# -------------------------------------------------------------------------------------------------


def _torch():
    import torch

    return torch


def device():
    return _torch().device("cuda", _torch().cuda.current_device())


def to_device(a, dtype=None):
    """numpy / torch array -> contiguous device tensor (float64 unless dtype given)."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a
    else:
        t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is None:
        dtype = torch.float64 if t.is_floating_point() else t.dtype
    return t.to(device=device(), dtype=dtype).contiguous()


def _is_f32(a):
    """True for an FP32 numpy array or torch tensor."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.dtype == torch.float32
    return isinstance(a, np.ndarray) and a.dtype == np.float32


def _dp(t):
    """Device pointer of a contiguous torch tensor on the GPU."""
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "expected a contiguous GPU tensor"
    return _vp(t.data_ptr())


def _stream():
    return _vp(_torch().cuda.current_stream().cuda_stream)


def _empty(shape, dtype=None):
    torch = _torch()
    return torch.empty(shape, dtype=dtype or torch.float64, device=device())


def _out(out, shape, dtype=None):
    """The caller's output tensor (checked) or a new one."""
    if out is None:
        return _empty(shape, dtype)
    torch = _torch()
    want = dtype or torch.float64
    assert out.is_cuda and out.is_contiguous(), "output must be a contiguous GPU tensor"
    assert tuple(out.shape) == tuple(shape) and out.dtype == want, f"output must be {want} of shape {tuple(shape)}"
    return out


def _inplace(ret, fn):
    """Run fn(device_tensor) so that it writes into ret: a contiguous FP64 GPU tensor is written
    directly; a numpy array goes through a device copy that is copied back into it."""
    torch = _torch()
    if isinstance(ret, torch.Tensor) and ret.is_cuda:
        fn(ret)
        return ret
    assert isinstance(ret, np.ndarray) and ret.dtype == np.float64, "ret: FP64 GPU tensor or numpy array"
    d = to_device(ret)
    fn(d)
    torch.cuda.current_stream().synchronize()
    ret[...] = d.cpu().numpy().reshape(ret.shape)
    return ret


def _np_of(a):
    """Host numpy view of a scalar, numpy array or torch tensor."""
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def _dtype_code(t):
    torch = _torch()
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise TypeError(f"unsupported dtype {t.dtype}")


# -------------------------------------------------------------------------------------------------
# Mesh (host, set-up): include/GooseFEM/Mesh*.h
# -------------------------------------------------------------------------------------------------


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _nodeset(nx, ny, nz, name):
    n = _i64(0)
    _call("gf_mesh_nodeset", _i64(nx), _i64(ny), _i64(nz), name.encode(), None, ctypes.byref(n))
    out = np.empty(n.value, dtype=np.int64)
    _call("gf_mesh_nodeset", _i64(nx), _i64(ny), _i64(nz), name.encode(), _hp(out), ctypes.byref(n))
    return out


class _RegularBase:
    _nz = 0

    def __getattr__(self, name):
        if name.startswith("nodes"):
            r = _nodeset(self.nelx, self.nely, self._nz, name)
            return int(r[0]) if name.endswith("Corner") or name == "nodesOrigin" else r
        raise AttributeError(name)

    @property
    def dofs(self):
        """Mesh::dofs (Mesh.h:93-96)."""
        out = np.empty((self.nnode, self.ndim), dtype=np.int64)
        _call("gf_mesh_dofs", _i64(self.nnode), _i32(self.ndim), _hp(out))
        return out

    @property
    def nodesPeriodic(self):
        """RegularBase::nodesPeriodic (Mesh.h:511-540, 1577-1692)."""
        n = _i64(0)
        _call("gf_mesh_nodes_periodic", _i64(self.nelx), _i64(self.nely), _i64(self._nz), None, ctypes.byref(n))
        out = np.empty((n.value, 2), dtype=np.int64)
        _call("gf_mesh_nodes_periodic", _i64(self.nelx), _i64(self.nely), _i64(self._nz), _hp(out), ctypes.byref(n))
        return out

    @property
    def dofsPeriodic(self):
        """RegularBase::dofsPeriodic (Mesh.h:291-304)."""
        ties = self.nodesPeriodic
        out = np.empty((self.nnode, self.ndim), dtype=np.int64)
        _call("gf_mesh_dofs_periodic", _i64(self.nnode), _i32(self.ndim), _hp(ties), _i64(ties.shape[0]), _hp(out))
        return out


class _Quad4FineLayer:
    """Mesh::Quad4::FineLayer (include/GooseFEM/MeshQuad4.h:200-1135): fine element layers around
    the middle, coarsened by 3 in x towards the bottom and top edges (host set-up)."""

    def __init__(self, nelx, nely, h=1.0, nfine=1):
        self._args = (_i64(int(nelx)), _i64(int(nely)), _f64(float(h)), _i64(int(nfine)))
        self.h = float(h)
        self.nne, self.ndim = 4, 2
        sizes = np.zeros(4, dtype=np.int64)
        _call("gf_mesh_quad4_finelayer", *self._args, _hp(sizes), None, None)
        self.nelem, self.nnode, self.nelx, self.nely = (int(v) for v in sizes)

    def _set(self, name):
        n = _i64(0)
        _call("gf_mesh_quad4_finelayer_set", *self._args, name.encode(), None, ctypes.byref(n))
        out = np.empty(n.value, dtype=np.int64)
        _call("gf_mesh_quad4_finelayer_set", *self._args, name.encode(), _hp(out), ctypes.byref(n))
        return out

    def __getattr__(self, name):
        if name.startswith("nodes") or name.startswith("elements"):
            r = self._set(name)
            return int(r[0]) if name.endswith("Corner") or name == "nodesOrigin" else r
        raise AttributeError(name)

    @property
    def coor(self):
        out = np.empty((self.nnode, 2))
        _call("gf_mesh_quad4_finelayer", *self._args, None, _hp(out), None)
        return out

    @property
    def conn(self):
        out = np.empty((self.nelem, 4), dtype=np.int64)
        _call("gf_mesh_quad4_finelayer", *self._args, None, None, _hp(out))
        return out

    @property
    def dofs(self):
        out = np.empty((self.nnode, self.ndim), dtype=np.int64)
        _call("gf_mesh_dofs", _i64(self.nnode), _i32(self.ndim), _hp(out))
        return out

    @property
    def nodesPeriodic(self):
        return self._set("nodesPeriodic").reshape(-1, 2)

    @property
    def dofsPeriodic(self):
        ties = self.nodesPeriodic
        out = np.empty((self.nnode, self.ndim), dtype=np.int64)
        _call("gf_mesh_dofs_periodic", _i64(self.nnode), _i32(self.ndim), _hp(ties), _i64(ties.shape[0]), _hp(out))
        return out

    def elementsMiddleLayer(self):
        return self._set("elementsMiddleLayer")

    def roll(self, n):
        """FineLayer::roll (MeshQuad4.h:539-564): element map shifting every layer by n coarse
        elements in x; new_elemvar = elemvar[roll(n)]."""
        out = np.empty(self.nelem, dtype=np.int64)
        _call("gf_mesh_quad4_finelayer_roll", *self._args, _i64(int(n)), _hp(out))
        return out


class _Hex8FineLayer(_Quad4FineLayer):
    """Mesh::Hex8::FineLayer (include/GooseFEM/MeshHex8.h:613-2296): fine element layers around
    the middle, coarsened by 3 in x or z towards the bottom and top faces (host set-up)."""

    def __init__(self, nelx, nely, nelz, h=1.0, nfine=1):
        self._args = (_i64(int(nelx)), _i64(int(nely)), _i64(int(nelz)), _f64(float(h)), _i64(int(nfine)))
        self.h = float(h)
        self.nne, self.ndim = 8, 3
        sizes = np.zeros(5, dtype=np.int64)
        _call("gf_mesh_hex8_finelayer", *self._args, _hp(sizes), None, None)
        self.nelem, self.nnode, self.nelx, self.nely, self.nelz = (int(v) for v in sizes)

    def _set(self, name):
        n = _i64(0)
        _call("gf_mesh_hex8_finelayer_set", *self._args, name.encode(), None, ctypes.byref(n))
        out = np.empty(n.value, dtype=np.int64)
        _call("gf_mesh_hex8_finelayer_set", *self._args, name.encode(), _hp(out), ctypes.byref(n))
        return out

    def roll(self, n):
        """The reference's Hex8 FineLayer has no roll (only Mesh::Quad4::FineLayer::roll,
        MeshQuad4.h:539-564, exists)."""
        raise NotImplementedError("Mesh.Hex8.FineLayer has no roll (the reference defines it for Quad4 only)")

    @property
    def coor(self):
        out = np.empty((self.nnode, 3))
        _call("gf_mesh_hex8_finelayer", *self._args, None, _hp(out), None)
        return out

    @property
    def conn(self):
        out = np.empty((self.nelem, 8), dtype=np.int64)
        _call("gf_mesh_hex8_finelayer", *self._args, None, None, _hp(out))
        return out


class _Quad4Regular(_RegularBase):
    """Mesh::Quad4::Regular (include/GooseFEM/MeshQuad4.h:32-195)."""

    def __init__(self, nelx, nely, h=1.0):
        self.nelx, self.nely, self.h = int(nelx), int(nely), float(h)
        self.nne, self.ndim = 4, 2
        self.nelem = self.nelx * self.nely
        self.nnode = (self.nelx + 1) * (self.nely + 1)

    @property
    def coor(self):
        out = np.empty((self.nnode, 2))
        _call("gf_mesh_quad4_regular", _i64(self.nelx), _i64(self.nely), _f64(self.h), _hp(out), None)
        return out

    @property
    def conn(self):
        out = np.empty((self.nelem, 4), dtype=np.int64)
        _call("gf_mesh_quad4_regular", _i64(self.nelx), _i64(self.nely), _f64(self.h), None, _hp(out))
        return out

    @property
    def elementgrid(self):
        return np.arange(self.nelem).reshape(self.nely, self.nelx)


class _Hex8Regular(_RegularBase):
    """Mesh::Hex8::Regular (include/GooseFEM/MeshHex8.h:26-609)."""

    def __init__(self, nelx, nely, nelz, h=1.0):
        self.nelx, self.nely, self._nz, self.h = int(nelx), int(nely), int(nelz), float(h)
        self.nelz = self._nz
        self.nne, self.ndim = 8, 3
        self.nelem = self.nelx * self.nely * self.nelz
        self.nnode = (self.nelx + 1) * (self.nely + 1) * (self.nelz + 1)

    @property
    def coor(self):
        out = np.empty((self.nnode, 3))
        _call("gf_mesh_hex8_regular", _i64(self.nelx), _i64(self.nely), _i64(self.nelz), _f64(self.h), _hp(out), None)
        return out

    @property
    def conn(self):
        out = np.empty((self.nelem, 8), dtype=np.int64)
        _call("gf_mesh_hex8_regular", _i64(self.nelx), _i64(self.nely), _i64(self.nelz), _f64(self.h), None, _hp(out))
        return out


def _csr_lists(ptr, idx):
    return [list(map(int, idx[ptr[i] : ptr[i + 1]])) for i in range(len(ptr) - 1)]


def _renumber(a):
    a = _i64a(a)
    out = np.empty_like(a)
    _call("gf_mesh_renumber", _hp(a), _i64(a.size), _hp(out))
    return out


def _coordination(conn):
    conn = _i64a(conn)
    nv = int(conn.max()) + 1
    out = np.empty(nv, dtype=np.int64)
    _call("gf_mesh_coordination", _hp(conn), _i64(conn.size), _i64(nv), _hp(out))
    return out


def _elem2node(conn, dofs=None):
    conn = _i64a(conn)
    nelem, nne = conn.shape
    if dofs is None:
        nnode, ndim, dp = int(conn.max()) + 1, 1, None
    else:
        dofs = _i64a(dofs)
        (nnode, ndim), dp = dofs.shape, _hp(dofs)
    nnz = _i64(0)
    ptr = np.empty(nnode + 1, dtype=np.int64)
    _call("gf_mesh_elem2node", _hp(conn), _i64(nelem), _i32(nne), dp, _i64(nnode), _i32(ndim), _hp(ptr), None,
          ctypes.byref(nnz))
    idx = np.empty(nnz.value, dtype=np.int64)
    _call("gf_mesh_elem2node", _hp(conn), _i64(nelem), _i32(nne), dp, _i64(nnode), _i32(ndim), _hp(ptr), _hp(idx),
          ctypes.byref(nnz))
    return _csr_lists(ptr, idx)


def _nodaltyings(dofs):
    dofs = _i64a(dofs)
    nnode, ndim = dofs.shape
    nnz = _i64(0)
    ptr = np.empty(nnode + 1, dtype=np.int64)
    _call("gf_mesh_nodaltyings", _hp(dofs), _i64(nnode), _i32(ndim), _hp(ptr), None, ctypes.byref(nnz))
    idx = np.empty(nnz.value, dtype=np.int64)
    _call("gf_mesh_nodaltyings", _hp(dofs), _i64(nnode), _i32(ndim), _hp(ptr), _hp(idx), ctypes.byref(nnz))
    return _csr_lists(ptr, idx)


def _quad4_only(coor, conn):
    coor, conn = np.asarray(coor, dtype=np.float64), _i64a(conn)
    if coor.ndim != 2 or conn.ndim != 2 or coor.shape[1] != 2 or conn.shape[1] != 4:
        raise NotImplementedError("Element-type not implemented")
    if conn.size and conn.max() >= coor.shape[0]:
        raise ValueError("conn refers to a node beyond coor")
    return coor, conn


def _edgesize(coor, conn, elementType=None):
    """Mesh::edgesize (Mesh.h:2571-2615): the length of each Quad4 element edge, [nelem, 4],
    edge i running from local node i to local node i + 1 (host set-up)."""
    coor, conn = _quad4_only(coor, conn)
    x, y = coor[conn, 0], coor[conn, 1]
    dx, dy = np.roll(x, -1, axis=1) - x, np.roll(y, -1, axis=1) - y
    return np.sqrt(dx * dx + dy * dy)


def _centers(coor, conn, elementType=None):
    """Mesh::centers (Mesh.h:2626-2659): the mean of each Quad4 element's nodal coordinates, the
    four nodes summed in local order, then divided by 4 (host set-up)."""
    coor, conn = _quad4_only(coor, conn)
    ret = np.zeros((conn.shape[0], coor.shape[1]))
    for i in range(4):
        ret += coor[conn[:, i]]
    return ret / 4.0


def _elemmap2nodemap(elem_map, coor, conn, elementType=None):
    """Mesh::elemmap2nodemap (Mesh.h:2670-2709): the node map that goes with an element map
    (new_elemvar = elemvar[elem_map]) on a Quad4 mesh.  Local node column i = 0..3 maps the nodes
    in that column of conn to the same column of conn[elem_map] (the last element that lists a
    node wins); a node keeps the first column's mapping that reaches it (host set-up)."""
    coor, conn = _quad4_only(coor, conn)
    elem_map = _i64a(elem_map)
    if elem_map.shape != (conn.shape[0],):
        raise ValueError("elem_map must have one entry per element")
    N = coor.shape[0]
    ret = np.full(N, N, dtype=np.int64)
    for i in range(4):
        t = np.full(N, N, dtype=np.int64)
        t[conn[:, i]] = conn[elem_map, i]
        ret = np.where(ret == N, t, ret)
    return ret


class _Tri3Regular(_Quad4Regular):
    """Mesh::Tri3::Regular (include/GooseFEM/MeshTri3.h:26-232): the Quad4 Regular grid's nodes,
    node sets and periodic pairs, each square split into two triangles (host set-up)."""

    def __init__(self, nelx, nely, h=1.0):
        super().__init__(nelx, nely, h)
        self.nne = 3
        self.nelem = 2 * self.nelx * self.nely

    @property
    def conn(self):
        nx = self.nelx
        iy, ix = np.meshgrid(np.arange(self.nely), np.arange(nx), indexing="ij")
        n0 = (iy * (nx + 1) + ix).ravel()
        out = np.empty((self.nelem, 3), dtype=np.int64)
        out[0::2] = np.stack((n0, n0 + 1, n0 + nx + 1), axis=1)
        out[1::2] = np.stack((n0 + 1, n0 + nx + 2, n0 + nx + 1), axis=1)
        return out


def _tri3_get_orientation(coor, conn):
    """Mesh::Tri3::getOrientation (MeshTri3.h:240-268): -1 for clockwise elements, +1 otherwise."""
    coor, conn = np.asarray(coor, dtype=np.float64), _i64a(conn)
    v1 = coor[conn[:, 0]] - coor[conn[:, 1]]
    v2 = coor[conn[:, 2]] - coor[conn[:, 1]]
    k = v1[:, 0] * v2[:, 1] - v2[:, 0] * v1[:, 1]
    return np.where(k < 0, -1, 1).astype(np.int32)


def _tri3_set_orientation(coor, conn, *args):
    """Mesh::Tri3::setOrientation(coor, conn[, val], orientation=-1) (MeshTri3.h:279-326): swap the
    last two nodes of the elements whose orientation differs from the requested one."""
    if len(args) == 2:
        val, orientation = args
    else:
        orientation = args[0] if args else -1
        val = _tri3_get_orientation(coor, conn)
    assert orientation in (-1, 1)
    out = _i64a(conn).copy()
    flip = (np.asarray(val) > 0) if orientation == -1 else (np.asarray(val) < 0)
    out[flip, 1], out[flip, 2] = out[flip, 2].copy(), out[flip, 1].copy()
    return out


def _nodal_mass(coor, conn):
    """Mesh::nodal_mass (Mesh.h:2732-2752): the lumped (nodal-quadrature) mass of unit density per
    node and direction, Quad4 only (the reference's only implemented type), on the GPU."""
    coor = np.asarray(coor, dtype=np.float64)
    conn = _i64a(conn)
    assert conn.shape[1] == 4 and coor.shape[1] == 2, "nodal_mass: Quad4 meshes only"
    dofs = np.arange(coor.size, dtype=np.int64).reshape(coor.shape)
    vec = Vector(conn, dofs)
    M = MatrixDiagonal(conn, dofs)
    quad = Element.Quad4.Quadrature(vec.AsElement(coor), Element.Quad4.Nodal.xi(), Element.Quad4.Nodal.w())
    M.assemble(quad.Int_N_scalar_NT_dV(np.ones(conn.shape)))
    return vec.AsNode(M.data)


def _center_of_gravity(coor, conn):
    """Mesh::center_of_gravity (Mesh.h:2809-2831): nodal-mass weighted mean of the coordinates."""
    m = _nodal_mass(coor, conn)
    c = to_device(coor)
    return (c * m).sum(0) / m.sum(0)


class _StopList:
    """Iterate::StopList (include/GooseFEM/Iterate.h): the last n residuals of an iteration, for a
    stopping test (all below a tolerance, and optionally descending)."""

    def __init__(self, n=1):
        self.reset(n)

    def reset(self, n=None):
        if n is not None:
            self._n = int(n)
        self._res = [float("inf")] * self._n

    def roll_insert(self, res):
        self._res = self._res[1:] + [float(res)]

    def descending(self):
        return all(a >= b for a, b in zip(self._res, self._res[1:]))

    def all_less(self, tol):
        return not any(r >= tol for r in self._res)

    @property
    def data(self):
        return list(self._res)


Iterate = SimpleNamespace(StopList=_StopList)


Mesh = SimpleNamespace(
    Quad4=SimpleNamespace(Regular=_Quad4Regular, FineLayer=_Quad4FineLayer),
    Hex8=SimpleNamespace(Regular=_Hex8Regular, FineLayer=_Hex8FineLayer),
    Tri3=SimpleNamespace(Regular=_Tri3Regular, getOrientation=_tri3_get_orientation,
                         setOrientation=_tri3_set_orientation),
    dofs=lambda nnode, ndim: np.arange(nnode * ndim, dtype=np.int64).reshape(nnode, ndim),
    renumber=_renumber,
    nodal_mass=_nodal_mass,
    center_of_gravity=_center_of_gravity,
    centers=_centers,
    edgesize=_edgesize,
    elemmap2nodemap=_elemmap2nodemap,
    coordination=_coordination,
    elem2node=_elem2node,
    nodaltyings=_nodaltyings,
    node2dof=lambda dofs: _elem2node(dofs),
)


# -------------------------------------------------------------------------------------------------
# Tyings (host, set-up): include/GooseFEM/TyingsPeriodic.h
# -------------------------------------------------------------------------------------------------


class _Control:
    """Tyings::Control (TyingsPeriodic.h:361-439)."""

    def __init__(self, coor, dofs):
        coor = np.ascontiguousarray(coor, dtype=np.float64)
        dofs = _i64a(dofs)
        nnode, ndim = coor.shape
        self.coor = np.empty((nnode + ndim, ndim))
        self.dofs = np.empty((nnode + ndim, ndim), dtype=np.int64)
        self.controlDofs = np.empty((ndim, ndim), dtype=np.int64)
        self.controlNodes = np.empty(ndim, dtype=np.int64)
        _call("gf_tyings_control", _hp(coor), _hp(dofs), _i64(nnode), _i32(ndim), _hp(self.coor), _hp(self.dofs),
              _hp(self.controlDofs), _hp(self.controlNodes))


class _Periodic:
    """Tyings::Periodic (TyingsPeriodic.h:82-343).  Tying matrices are returned as scipy-free
    CSR triples (rowptr, col, val) via Cdi_csr / Cdu_csr / Cdp_csr, and dense via Cdi / Cdu / Cdp."""

    def __init__(self, coor, dofs, control_dofs, nodal_tyings, iip=()):
        self._coor = np.ascontiguousarray(coor, dtype=np.float64)
        dofs = _i64a(dofs)
        control_dofs = _i64a(control_dofs)
        self._ties = _i64a(nodal_tyings)
        iip = _i64a(iip)
        nnode, ndim = dofs.shape
        self.ndim = ndim
        self.dofs = np.empty_like(dofs)
        self.control = np.empty_like(control_dofs)
        sizes = np.empty(4, dtype=np.int64)
        _call("gf_tyings_periodic", _hp(dofs), _i64(nnode), _i32(ndim), _hp(control_dofs), _hp(self._ties),
              _i64(self._ties.shape[0]), _hp(iip), _i64(iip.size), _hp(self.dofs), _hp(self.control), _hp(sizes))
        self.nnu, self.nnp, self.nni, self.nnd = (int(s) for s in sizes)

    def _csr(self, which):
        nrows = self._ties.shape[0] * self.ndim
        rowptr = np.empty(nrows + 1, dtype=np.int64)
        nnz = _i64(0)
        args = (_i32(which), _hp(self._coor), _hp(self.dofs), _i32(self.ndim), _hp(self.control), _hp(self._ties),
                _i64(self._ties.shape[0]), _i64(self.nnu))
        _call("gf_tyings_matrix", *args, _hp(rowptr), None, None, ctypes.byref(nnz))
        col = np.empty(nnz.value, dtype=np.int64)
        val = np.empty(nnz.value)
        _call("gf_tyings_matrix", *args, _hp(rowptr), _hp(col), _hp(val), ctypes.byref(nnz))
        return rowptr, col, val

    def _dense(self, which, ncol):
        rowptr, col, val = self._csr(which)
        out = np.zeros((rowptr.size - 1, ncol))
        for r in range(rowptr.size - 1):
            out[r, col[rowptr[r]:rowptr[r + 1]]] += val[rowptr[r]:rowptr[r + 1]]
        return out

    # (rowptr, col, val, ncols): the form the tied classes take without densifying
    Cdi_csr = property(lambda self: self._csr(0) + (self.nni,))
    Cdu_csr = property(lambda self: self._csr(1) + (self.nnu,))
    Cdp_csr = property(lambda self: self._csr(2) + (self.nnp,))
    Cdi = property(lambda self: self._dense(0, self.nni))
    Cdu = property(lambda self: self._dense(1, self.nnu))
    Cdp = property(lambda self: self._dense(2, self.nnp))


Tyings = SimpleNamespace(Control=_Control, Periodic=_Periodic)


def randn(seed: int, n: int, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
    """Portable SplitMix64 + Box-Muller normal samples (host)."""
    out = np.empty(n)
    _call("gf_randn", seed, n, mean, std, _hp(out))
    return out


# -------------------------------------------------------------------------------------------------
# Vector (include/GooseFEM/Vector.h)
# -------------------------------------------------------------------------------------------------


class _Handle:
    _destroy = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and self._destroy and _lib is not None:
            try:
                f = getattr(_lib, self._destroy)
                f(h)
            except Exception:
                pass
            self._h = None


class Vector(_Handle):
    """GooseFEM::Vector (Vector.h:33-47): conversions elemvec <-> nodevec <-> dofval on the GPU."""

    _destroy = "gf_vector_destroy"

    def __init__(self, conn, dofs):
        conn = _i64a(conn)
        dofs = _i64a(dofs)
        self.nelem, self.nne = conn.shape
        self.nnode, self.ndim = dofs.shape
        self.conn, self.dofs_h = conn, dofs
        h = _vp()
        _call("gf_vector_create", ctypes.byref(h), _hp(conn), _i64(self.nelem), _i32(self.nne), _hp(dofs),
              _i64(self.nnode), _i32(self.ndim))
        self._h = h
        nd = _i64(0)
        _call("gf_vector_sizes", h, None, None, None, None, ctypes.byref(nd))
        self.ndof = nd.value

    @property
    def dofs(self):
        return self.dofs_h

    def shape_elemvec(self):
        return (self.nelem, self.nne, self.ndim)

    def shape_nodevec(self):
        return (self.nnode, self.ndim)

    def shape_dofval(self):
        return (self.ndof,)

    def AsElement(self, arg, out=None):
        """AsElement(nodevec | dofval) (Vector.h:193-211)."""
        arg = to_device(arg)
        out = _out(out, self.shape_elemvec())
        if arg.dim() == 2:
            _call("gf_vector_as_element_nodevec", self._h, _dp(arg), _dp(out), _stream())
        else:
            assert arg.numel() == self.ndof
            _call("gf_vector_as_element_dofval", self._h, _dp(arg), _dp(out), _stream())
        return out

    def AsDofs(self, arg, out=None):
        """AsDofs(nodevec | elemvec) (Vector.h:141-159): overwrite, last in loop order wins."""
        arg = to_device(arg)
        out = _out(out, self.shape_dofval())
        if arg.dim() == 2:
            _call("gf_vector_as_dofs_nodevec", self._h, _dp(arg), _dp(out), _stream())
        else:
            _call("gf_vector_as_dofs_elemvec", self._h, _dp(arg), _dp(out), _stream())
        return out

    def AsNode(self, arg, out=None):
        """AsNode(dofval | elemvec) (Vector.h:167-185)."""
        arg = to_device(arg)
        out = _out(out, self.shape_nodevec())
        if arg.dim() == 1:
            _call("gf_vector_as_node_dofval", self._h, _dp(arg), _dp(out), _stream())
        else:
            _call("gf_vector_as_node_elemvec", self._h, _dp(arg), _dp(out), _stream())
        return out

    def AssembleDofs(self, arg, out=None):
        """AssembleDofs(nodevec | elemvec) (Vector.h:219-237); elemvec may be float32."""
        torch = _torch()
        arg = arg if isinstance(arg, torch.Tensor) and arg.is_cuda else to_device(arg)
        out = _out(out, self.shape_dofval())
        if arg.dim() == 2:
            _call("gf_vector_assemble_dofs_nodevec", self._h, _dp(arg.double().contiguous()), _dp(out), _stream())
        else:
            arg = arg.contiguous()
            _call("gf_vector_assemble_dofs_elemvec", self._h, _dp(arg), _i32(_dtype_code(arg)), _dp(out), _stream())
        return out

    def AssembleNode(self, arg, out=None):
        """AssembleNode(elemvec) (Vector.h:245-263) = AsNode(AssembleDofs(elemvec))."""
        torch = _torch()
        arg = arg.contiguous() if isinstance(arg, torch.Tensor) and arg.is_cuda else to_device(arg)
        out = _out(out, self.shape_nodevec())
        _call("gf_vector_assemble_node_elemvec", self._h, _dp(arg), _i32(_dtype_code(arg)), _dp(out), _stream())
        return out

    # write-into forms of the reference (Vector.h: asElement, asDofs, asNode, assembleDofs,
    # assembleNode): ret is a contiguous FP64 GPU tensor, or a numpy array updated in place
    def asElement(self, arg, ret):
        return _inplace(ret, lambda r: self.AsElement(arg, out=r))

    def asDofs(self, arg, ret):
        return _inplace(ret, lambda r: self.AsDofs(arg, out=r))

    def asNode(self, arg, ret):
        return _inplace(ret, lambda r: self.AsNode(arg, out=r))

    def assembleDofs(self, arg, ret):
        return _inplace(ret, lambda r: self.AssembleDofs(arg, out=r))

    def assembleNode(self, arg, ret):
        return _inplace(ret, lambda r: self.AssembleNode(arg, out=r))


class VectorPartitioned(Vector):
    """GooseFEM::VectorPartitioned (VectorPartitioned.h:72-661): a Vector whose DOFs split into
    unknown (iiu, ascending) and prescribed (iip) ones; all conversions run on the GPU."""

    def __init__(self, conn, dofs, iip):
        super().__init__(conn, dofs)
        iip = _i64a(iip).ravel()
        h = _vp()
        _call("gf_vecpart_create", ctypes.byref(h), self._h, _hp(iip), _i64(iip.size))
        self._ph = h
        nu, npp = _i64(0), _i64(0)
        _call("gf_vecpart_sizes", h, ctypes.byref(nu), ctypes.byref(npp))
        self.nnu, self.nnp = nu.value, npp.value
        self._iiu = np.empty(self.nnu, dtype=np.int64)
        self._iip = np.empty(self.nnp, dtype=np.int64)
        self._part = np.empty((self.nnode, self.ndim), dtype=np.int64)
        _call("gf_vecpart_index", h, _hp(self._iiu), _hp(self._iip), _hp(self._part))

    def __del__(self):
        h = getattr(self, "_ph", None)
        if h is not None and _lib is not None:
            try:
                _lib.gf_vecpart_destroy(h)
            except Exception:
                pass
            self._ph = None
        super().__del__()

    @property
    def iiu(self):
        return self._iiu.copy()

    @property
    def iip(self):
        return self._iip.copy()

    def dofs_is_u(self):
        return self._part < self.nnu

    def dofs_is_p(self):
        return self._part >= self.nnu

    @staticmethod
    def _kind(t, ndof):
        if t.dim() == 1:
            assert t.numel() == ndof
            return 0
        return 1 if t.dim() == 2 else 2

    def AsDofs_u(self, arg):
        """AsDofs_u(dofval | nodevec | elemvec) (VectorPartitioned.h:428-540)."""
        arg = to_device(arg)
        out = _empty((self.nnu,))
        _call("gf_vecpart_as_dofs", self._ph, _i32(self._kind(arg, self.ndof)), _dp(arg), _dp(out), None, _stream())
        return out

    def AsDofs_p(self, arg):
        """AsDofs_p(dofval | nodevec | elemvec) (VectorPartitioned.h:549-661)."""
        arg = to_device(arg)
        out = _empty((self.nnp,))
        _call("gf_vecpart_as_dofs", self._ph, _i32(self._kind(arg, self.ndof)), _dp(arg), None, _dp(out), _stream())
        return out

    def _from(self, kind, shape, dofval_u, dofval_p):
        u, p = to_device(dofval_u), to_device(dofval_p)
        assert u.numel() == self.nnu and p.numel() == self.nnp
        out = _empty(shape)
        _call("gf_vecpart_from_partitioned", self._ph, _i32(kind), _dp(u), _dp(p), _dp(out), _stream())
        return out

    def DofsFromParitioned(self, dofval_u, dofval_p):
        """DofsFromParitioned (sic, VectorPartitioned.h:276-307)."""
        return self._from(0, self.shape_dofval(), dofval_u, dofval_p)

    def NodeFromPartitioned(self, dofval_u, dofval_p):
        """NodeFromPartitioned (VectorPartitioned.h:325-358)."""
        return self._from(1, self.shape_nodevec(), dofval_u, dofval_p)

    def ElementFromPartitioned(self, dofval_u, dofval_p):
        """ElementFromPartitioned (VectorPartitioned.h:376-412)."""
        return self._from(2, self.shape_elemvec(), dofval_u, dofval_p)

    def _copy(self, which, src, dest):
        src = to_device(src)
        _call("gf_vecpart_copy", self._ph, _i32(which), _dp(src), _dp(dest), _stream())
        return dest

    def Copy_u(self, nodevec_src, nodevec_dest):
        """Copy_u (VectorPartitioned.h:176-212): nodevec_dest with its unknown entries from src."""
        return self._copy(0, nodevec_src, to_device(nodevec_dest).clone())

    def Copy_p(self, nodevec_src, nodevec_dest):
        """Copy_p (VectorPartitioned.h:229-263): nodevec_dest with its prescribed entries from src."""
        return self._copy(1, nodevec_src, to_device(nodevec_dest).clone())

    def copy_u(self, nodevec_src, nodevec_dest):
        """copy_u: in place on nodevec_dest (FP64 GPU tensor or numpy array)."""
        return _inplace(nodevec_dest, lambda r: self._copy(0, nodevec_src, r))

    def copy_p(self, nodevec_src, nodevec_dest):
        """copy_p: in place on nodevec_dest (FP64 GPU tensor or numpy array)."""
        return _inplace(nodevec_dest, lambda r: self._copy(1, nodevec_src, r))


# -------------------------------------------------------------------------------------------------
# Quadrature (include/GooseFEM/Element*.h)
# -------------------------------------------------------------------------------------------------


class _QuadShapes:
    """Shapes, allocation and AsTensor of QuadratureBase (Element.h:201-420), on the GPU."""

    def shape_elemvec(self, arg=None):
        return (self.nelem, self.nne, self.ndim if arg is None else arg)

    def shape_elemmat(self):
        return (self.nelem, self.nne * self.ndim, self.nne * self.ndim)

    def shape_qscalar(self):
        return (self.nelem, self.nip)

    def shape_qvector(self, arg=None):
        return (self.nelem, self.nip, self.tdim if arg is None else arg)

    def shape_qtensor(self, rank=2, arg=None):
        return (self.nelem, self.nip) + ((self.tdim if arg is None else arg),) * rank

    def _alloc(self, shape, val):
        torch = _torch()
        if val is None:
            return _empty(shape)
        return torch.full(shape, float(val), dtype=torch.float64, device=device())

    def allocate_elemvec(self, val=None):
        return self._alloc(self.shape_elemvec(), val)

    def allocate_elemmat(self, val=None):
        return self._alloc(self.shape_elemmat(), val)

    def allocate_qscalar(self, val=None):
        return self._alloc(self.shape_qscalar(), val)

    def allocate_qtensor(self, rank=2, val=None):
        return self._alloc(self.shape_qtensor(rank), val)

    def AsTensor(self, rank, qscalar):
        """AsTensor(rank, qscalar) (Element.h:234-245, Allocate.h:158-170): every component of the
        rank-`rank` tensor at a point equals that point's scalar."""
        qs = to_device(qscalar)
        assert tuple(qs.shape) == self.shape_qscalar()
        return qs.reshape(qs.shape + (1,) * rank).expand(self.shape_qtensor(rank)).contiguous()

    def asTensor(self, qscalar, qtensor):
        rank = qtensor.dim() - 2
        return _inplace(qtensor, lambda r: r.copy_(self.AsTensor(rank, qscalar)))


class _Quadrature(_QuadShapes, _Handle):
    _destroy = "gf_quad_destroy"
    _etype = None

    def __init__(self, x, xi=None, w=None, thick=1.0, rule=GAUSS):
        x = to_device(x)
        self.nelem = x.shape[0]
        xi_p = w_p = None
        nip = 0
        if xi is not None:
            self._xi = np.ascontiguousarray(xi, dtype=np.float64)
            self._w = np.ascontiguousarray(w, dtype=np.float64)
            xi_p, w_p, nip = _hp(self._xi), _hp(self._w), self._w.size
        h = _vp()
        _call("gf_quad_create", ctypes.byref(h), _i32(self._etype), _dp(x), _i64(self.nelem), _i32(rule), xi_p, w_p,
              _i32(nip), _f64(thick), _stream())
        self._h = h
        ne, nne, nd, td, ni = _i64(), _i32(), _i32(), _i32(), _i32()
        _call("gf_quad_sizes", h, ctypes.byref(ne), ctypes.byref(nne), ctypes.byref(nd), ctypes.byref(td),
              ctypes.byref(ni))
        self.nne, self.ndim, self.tdim, self.nip = nne.value, nd.value, td.value, ni.value

    def _wrap(self, ptr, shape):
        # copy out of the internal buffer (a view would outlive the handle)
        out = _empty(shape)
        _hip_memcpy_d2d(out, ptr, int(np.prod(shape)) * 8)
        return out

    @property
    def GradN(self):
        """GradN() (Element.h:575-578) [nelem, nip, nne, ndim] (copy of the device array)."""
        p = _vp()
        _call("gf_quad_gradN", self._h, ctypes.byref(p))
        return self._wrap(p, (self.nelem, self.nip, self.nne, self.ndim))

    @property
    def dV(self):
        """dV() (Element.h:584-587) [nelem, nip]."""
        p = _vp()
        _call("gf_quad_dV", self._h, ctypes.byref(p))
        return self._wrap(p, (self.nelem, self.nip))

    def update_x(self, x):
        xd = to_device(x)  # held until the call has enqueued its kernels
        _call("gf_quad_update_x", self._h, _dp(xd), _stream())

    def _grad(self, fn, elemvec, out=None):
        torch = _torch()
        ev = elemvec.contiguous() if isinstance(elemvec, torch.Tensor) and elemvec.is_cuda else to_device(elemvec)
        out = _out(out, self.shape_qtensor(2), dtype=ev.dtype)
        _call(fn, self._h, _dp(ev), _dp(out), _i32(_dtype_code(ev)), _stream())
        return out

    def GradN_vector(self, elemvec, out=None):
        return self._grad("gf_quad_gradN_vector", elemvec, out)

    def GradN_vector_T(self, elemvec, out=None):
        return self._grad("gf_quad_gradN_vector_T", elemvec, out)

    def SymGradN_vector(self, elemvec, out=None):
        return self._grad("gf_quad_symgradN_vector", elemvec, out)

    def InterpQuad_vector(self, elemvec, out=None):
        ev = to_device(elemvec)
        n = ev.shape[2]
        out = _out(out, (self.nelem, self.nip, n))
        _call("gf_quad_interp_quad_vector", self._h, _dp(ev), _i32(n), _dp(out), _stream())
        return out

    def Int_N_vector_dV(self, qvector, out=None):
        qv = to_device(qvector)
        n = qv.shape[2]
        out = _out(out, (self.nelem, self.nne, n))
        _call("gf_quad_int_N_vector_dV", self._h, _dp(qv), _i32(n), _dp(out), _stream())
        return out

    def Int_N_scalar_NT_dV(self, qscalar, out=None):
        qs = to_device(qscalar)
        K = self.nne * self.ndim
        out = _out(out, (self.nelem, K, K))
        _call("gf_quad_int_N_scalar_NT_dV", self._h, _dp(qs), _dp(out), _stream())
        return out

    def interpQuad_vector(self, elemvec, qvector):
        return _inplace(qvector, lambda r: self.InterpQuad_vector(elemvec, out=r))

    def int_N_vector_dV(self, qvector, elemvec):
        return _inplace(elemvec, lambda r: self.Int_N_vector_dV(qvector, out=r))

    def int_N_scalar_NT_dV(self, qscalar, elemmat):
        return _inplace(elemmat, lambda r: self.Int_N_scalar_NT_dV(qscalar, out=r))

    def Int_gradN_dot_tensor2_dV(self, qtensor, out=None):
        torch = _torch()
        qt = qtensor.contiguous() if isinstance(qtensor, torch.Tensor) and qtensor.is_cuda else to_device(qtensor)
        out = _out(out, (self.nelem, self.nne, self.ndim), dtype=qt.dtype)
        _call("gf_quad_int_gradN_dot_tensor2_dV", self._h, _dp(qt), _dp(out), _i32(_dtype_code(qt)), _stream())
        return out

    def Int_gradN_dot_tensor4_dot_gradNT_dV(self, qtensor, precision=F64, out=None):
        """Element stiffness.  With precision BF16X3 an FP32 qtensor gives an FP32 elemmat (the MFMA
        kernel's operands and accumulation are FP32); otherwise input and output are FP64."""
        torch = _torch()
        K = self.nne * self.ndim
        if precision == BF16X3 and _is_f32(qtensor):
            qt = to_device(qtensor, torch.float32)
            if out is None:
                out = _empty((self.nelem, K, K), torch.float32)
            _call("gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV_f32", self._h, _dp(qt), _dp(out), _stream())
            return out
        qt = to_device(qtensor)
        if out is None:
            out = _empty((self.nelem, K, K))
        _call("gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV", self._h, _dp(qt), _dp(out), _i32(precision), _stream())
        return out

    # write-into forms of the reference (Element.h: gradN_vector, gradN_vector_T, symGradN_vector,
    # int_gradN_dot_tensor2_dV, int_gradN_dot_tensor4_dot_gradNT_dV)
    def gradN_vector(self, elemvec, qtensor):
        return _inplace(qtensor, lambda r: self.GradN_vector(elemvec, out=r))

    def gradN_vector_T(self, elemvec, qtensor):
        return _inplace(qtensor, lambda r: self.GradN_vector_T(elemvec, out=r))

    def symGradN_vector(self, elemvec, qtensor):
        return _inplace(qtensor, lambda r: self.SymGradN_vector(elemvec, out=r))

    def int_gradN_dot_tensor2_dV(self, qtensor, elemvec):
        return _inplace(elemvec, lambda r: self.Int_gradN_dot_tensor2_dV(qtensor, out=r))

    def int_gradN_dot_tensor4_dot_gradNT_dV(self, qtensor, elemmat):
        return _inplace(elemmat, lambda r: self.Int_gradN_dot_tensor4_dot_gradNT_dV(qtensor, out=r))


def _hip_memcpy_d2d(dst, src_ptr, nbytes):
    """Device-to-device copy from a library-owned buffer into a torch tensor (current stream)."""
    hip = _hip_runtime()
    rc = hip.hipMemcpyAsync(_vp(dst.data_ptr()), src_ptr, ctypes.c_size_t(nbytes), 3, _stream())
    if rc != 0:
        raise GooseFEMError(GF_EHIP, f"hipMemcpyAsync failed ({rc})")


_hip = None


def _hip_runtime():
    global _hip
    if _hip is None:
        lib()
        _hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime already mapped (same SONAME)
        _hip.hipMemcpyAsync.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]
        _hip.hipMemcpyAsync.restype = ctypes.c_int
    return _hip


class _Quad4Quadrature(_Quadrature):
    """Element::Quad4::Quadrature (ElementQuad4.h:229-547), tdim = 2."""

    _etype = QUAD4

    def __init__(self, x, xi=None, w=None):
        super().__init__(x, xi, w)


class _Quad4QuadraturePlanar(_Quadrature):
    """Element::Quad4::QuadraturePlanar (ElementQuad4Planar.h:37-335), tdim = 3, plane strain."""

    _etype = QUAD4_PLANAR

    def __init__(self, x, xi=None, w=None, thick=1.0):
        super().__init__(x, xi, w, thick)


class _Hex8Quadrature(_Quadrature):
    """Element::Hex8::Quadrature (ElementHex8.h:206-400)."""

    _etype = HEX8

    def __init__(self, x, xi=None, w=None):
        super().__init__(x, xi, w)


class _Quad4QuadratureAxisymmetric(_QuadShapes, _Handle):
    """Element::Quad4::QuadratureAxisymmetric (ElementQuad4Axisymmetric.h:33-467): nodal coordinates
    (z, r), tensors 3 x 3 in (r, theta, z) order; all operators on the GPU, bit-exact in FP64."""

    _destroy = "gf_axi_destroy"

    def __init__(self, x, xi=None, w=None):
        if xi is None:
            xi, w = _rule(2, False)
        self._xi = np.ascontiguousarray(xi, dtype=np.float64)
        self._w = np.ascontiguousarray(w, dtype=np.float64)
        self._x = to_device(x)
        self.nelem, self.nne, self.ndim, self.tdim = int(self._x.shape[0]), 4, 2, 3
        self.nip = int(self._w.size)
        h = _vp()
        _call("gf_axi_create", ctypes.byref(h), _i64(self.nelem), _i32(self.nip), _dp(self._x),
              self._xi.ctypes.data_as(ctypes.c_void_p), self._w.ctypes.data_as(ctypes.c_void_p), _stream())
        self._h = h

    def update_x(self, x):
        self._x = to_device(x)
        _call("gf_axi_update_x", self._h, _dp(self._x), _stream())

    def _geometry(self):
        G, dV = _vp(), _vp()
        _call("gf_axi_geometry", self._h, ctypes.byref(G), ctypes.byref(dV))
        return G, dV

    @property
    def dV(self):
        """dV() [nelem, nip] (copy)."""
        out = _empty((self.nelem, self.nip))
        _hip_memcpy_d2d(out, self._geometry()[1], 8 * out.numel())
        return out

    def B(self):
        """B() [nelem, nip, nne, 3, 3, 3] (ElementQuad4Axisymmetric.h:401-438), built from the five
        non-zero components per node."""
        torch = _torch()
        G = _empty((self.nelem, self.nip, 4, 3))
        _hip_memcpy_d2d(G, self._geometry()[0], 8 * G.numel())
        B = torch.zeros((self.nelem, self.nip, 4, 3, 3, 3), dtype=torch.float64, device=device())
        B[..., 0, 0, 0] = G[..., 0]
        B[..., 0, 2, 2] = G[..., 0]
        B[..., 1, 1, 0] = G[..., 1]
        B[..., 2, 0, 0] = G[..., 2]
        B[..., 2, 2, 2] = G[..., 2]
        return B

    def _grad(self, mode, elemvec, out=None):
        ev = to_device(elemvec)
        out = _out(out, self.shape_qtensor(2))
        _call("gf_axi_grad", self._h, _i32(mode), _dp(ev), _dp(out), _stream())
        return out

    def GradN_vector(self, elemvec, out=None):
        return self._grad(0, elemvec, out)

    def GradN_vector_T(self, elemvec, out=None):
        return self._grad(1, elemvec, out)

    def SymGradN_vector(self, elemvec, out=None):
        return self._grad(2, elemvec, out)

    def Int_N_scalar_NT_dV(self, qscalar, out=None):
        qs = to_device(qscalar)
        out = _out(out, (self.nelem, 8, 8))
        _call("gf_axi_int_N_scalar_NT_dV", self._h, _dp(qs), _dp(out), _stream())
        return out

    def Int_gradN_dot_tensor2_dV(self, qtensor, out=None):
        qt = to_device(qtensor)
        out = _out(out, (self.nelem, 4, 2))
        _call("gf_axi_int_gradN_dot_tensor2_dV", self._h, _dp(qt), _dp(out), _stream())
        return out

    def Int_gradN_dot_tensor4_dot_gradNT_dV(self, qtensor, out=None):
        qt = to_device(qtensor)
        out = _out(out, (self.nelem, 8, 8))
        _call("gf_axi_int_gradN_dot_tensor4_dot_gradNT_dV", self._h, _dp(qt), _dp(out), _stream())
        return out

    def gradN_vector(self, elemvec, qtensor):
        return _inplace(qtensor, lambda r: self.GradN_vector(elemvec, out=r))

    def gradN_vector_T(self, elemvec, qtensor):
        return _inplace(qtensor, lambda r: self.GradN_vector_T(elemvec, out=r))

    def symGradN_vector(self, elemvec, qtensor):
        return _inplace(qtensor, lambda r: self.SymGradN_vector(elemvec, out=r))

    def int_gradN_dot_tensor2_dV(self, qtensor, elemvec):
        return _inplace(elemvec, lambda r: self.Int_gradN_dot_tensor2_dV(qtensor, out=r))

    def int_gradN_dot_tensor4_dot_gradNT_dV(self, qtensor, elemmat):
        return _inplace(elemmat, lambda r: self.Int_gradN_dot_tensor4_dot_gradNT_dV(qtensor, out=r))


def _rule(nd, nodal):
    a = 1.0 if nodal else 1.0 / np.sqrt(3.0)
    if nd == 2:
        s = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=np.float64)
    else:
        s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
    return s * a, np.ones(s.shape[0])


Element = SimpleNamespace(
    Quad4=SimpleNamespace(
        Quadrature=_Quad4Quadrature,
        QuadraturePlanar=_Quad4QuadraturePlanar,
        QuadratureAxisymmetric=_Quad4QuadratureAxisymmetric,
        Gauss=SimpleNamespace(xi=lambda: _rule(2, False)[0], w=lambda: _rule(2, False)[1], nip=lambda: 4),
        Nodal=SimpleNamespace(xi=lambda: _rule(2, True)[0], w=lambda: _rule(2, True)[1], nip=lambda: 4),
    ),
    Hex8=SimpleNamespace(
        Quadrature=_Hex8Quadrature,
        Gauss=SimpleNamespace(xi=lambda: _rule(3, False)[0], w=lambda: _rule(3, False)[1], nip=lambda: 8),
        Nodal=SimpleNamespace(xi=lambda: _rule(3, True)[0], w=lambda: _rule(3, True)[1], nip=lambda: 8),
    ),
)


def elemmat_dot(elemmat, elemvec):
    """out[e] = elemmat[e] . elemvec[e] for every element (the element part of sum_e K_e u_e); an
    FP32 elemmat is widened entry by entry, products and sums are FP64."""
    f32 = _is_f32(elemmat)
    Ke = to_device(elemmat, _torch().float32 if f32 else None)
    ue = to_device(elemvec)
    nelem, K = Ke.shape[0], Ke.shape[1]
    out = _empty(tuple(ue.shape))
    _call("gf_elemmat_dot_f32" if f32 else "gf_elemmat_dot", _i64(nelem), _i32(K), _dp(Ke), _dp(ue), _dp(out),
          _stream())
    return out


def elastic_stress(eps, K, G):
    """sigma = K tr(eps) I + 2 G dev(eps) per quadrature point (callers' GMat law)."""
    torch = _torch()
    eps = eps.contiguous() if isinstance(eps, torch.Tensor) and eps.is_cuda else to_device(eps)
    td = eps.shape[-1]
    nq = eps.numel() // (td * td)
    Kq = to_device(torch.as_tensor(K).expand(eps.shape[:-2]) if not isinstance(K, torch.Tensor)
                   else K.expand(eps.shape[:-2]), dtype=eps.dtype)
    Gq = to_device(torch.as_tensor(G).expand(eps.shape[:-2]) if not isinstance(G, torch.Tensor)
                   else G.expand(eps.shape[:-2]), dtype=eps.dtype)
    out = torch.empty_like(eps)
    _call("gf_material_elastic_stress", _i64(nq), _i32(td), _dp(eps), _dp(Kq), _dp(Gq), _dp(out),
          _i32(_dtype_code(eps)), _stream())
    return out


def elastic_tangent(K, G, td=3, dtype=None):
    """C_ijkl = K I(x)I + 2 G (I_sym - I(x)I / 3) per point, FP64 (default) or rounded to FP32
    (dtype torch.float32)."""
    torch = _torch()
    K = to_device(K)
    G = to_device(torch.as_tensor(G).expand(K.shape) if not isinstance(G, torch.Tensor) else G.expand(K.shape))
    f32 = dtype == torch.float32
    out = _empty(tuple(K.shape) + (td,) * 4, torch.float32 if f32 else None)
    _call("gf_material_elastic_tangent_f32" if f32 else "gf_material_elastic_tangent", _i64(K.numel()), _i32(td),
          _dp(K), _dp(G), _dp(out), _stream())
    return out


# -------------------------------------------------------------------------------------------------
# MatrixDiagonal / Matrix (include/GooseFEM/MatrixDiagonal.h, Matrix.h)
# -------------------------------------------------------------------------------------------------


class MatrixDiagonal(_Handle):
    """GooseFEM::MatrixDiagonal (MatrixDiagonal.h:132-280)."""

    _destroy = "gf_diag_destroy"

    def __init__(self, conn, dofs):
        self._vec = conn if isinstance(conn, Vector) else Vector(conn, dofs)
        h = _vp()
        _call("gf_diag_create", ctypes.byref(h), self._vec._h)
        self._h = h

    def assemble(self, elemmat):
        em = to_device(elemmat)
        _call("gf_diag_assemble", self._h, _dp(em), _stream())

    def set(self, A):
        Ad = to_device(A)
        _call("gf_diag_set", self._h, _dp(Ad), _stream())

    @property
    def data(self):
        p = _vp()
        _call("gf_diag_data", self._h, ctypes.byref(p))
        out = _empty((self._vec.ndof,))
        _hip_memcpy_d2d(out, p, self._vec.ndof * 8)
        return out

    def Dot(self, x):
        x = to_device(x)
        out = _empty(tuple(x.shape))
        fn = "gf_diag_dot_nodevec" if x.dim() == 2 else "gf_diag_dot_dofval"
        _call(fn, self._h, _dp(x), _dp(out), _stream())
        return out

    def Solve(self, b):
        b = to_device(b)
        out = _empty(tuple(b.shape))
        fn = "gf_diag_solve_nodevec" if b.dim() == 2 else "gf_diag_solve_dofval"
        _call(fn, self._h, _dp(b), _dp(out), _stream())
        return out

    def inverse_ptr(self):
        p = _vp()
        _call("gf_diag_inverse", self._h, ctypes.byref(p), _stream())
        return p


class MatrixDiagonalPartitioned(MatrixDiagonal):
    """GooseFEM::MatrixDiagonalPartitioned (MatrixDiagonalPartitioned.h:42-452).  assemble, set,
    data, Dot and Solve's factorisation act on the unpartitioned diagonal (MatrixDiagonal)."""

    def __init__(self, conn, dofs=None, iip=None):
        self._vec = conn if isinstance(conn, VectorPartitioned) else VectorPartitioned(conn, dofs, iip)
        h = _vp()
        _call("gf_diagpart_create", ctypes.byref(h), self._vec._ph)
        self._dph = h
        d = _vp()
        _call("gf_diagpart_diag", h, ctypes.byref(d))
        self._h = d  # the unpartitioned diagonal, owned by the partitioned matrix

    def __del__(self):
        h = getattr(self, "_dph", None)
        if h is not None and _lib is not None:
            try:
                _lib.gf_diagpart_destroy(h)
            except Exception:
                pass
        self._dph = None
        self._h = None

    @property
    def nnu(self):
        return self._vec.nnu

    @property
    def nnp(self):
        return self._vec.nnp

    def set(self, A):
        """set(A) (MatrixDiagonalPartitioned.h:121-135)."""
        Ad = to_device(A)
        _call("gf_diagpart_set", self._dph, _dp(Ad), _stream())

    @property
    def data_uu(self):
        out = _empty((self.nnu,))
        _call("gf_diagpart_data_part", self._dph, _dp(out), None, _stream())
        return out

    @property
    def data_pp(self):
        out = _empty((self.nnp,))
        _call("gf_diagpart_data_part", self._dph, None, _dp(out), _stream())
        return out

    def Dot_u(self, x_u, x_p):
        """Dot_u (MatrixDiagonalPartitioned.h:248-264): A_uu x_u."""
        x = to_device(x_u)
        out = _empty((self.nnu,))
        _call("gf_diagpart_dot_part", self._dph, _i32(0), _dp(x), _dp(out), _stream())
        return out

    def Dot_p(self, x_u, x_p):
        """Dot_p (MatrixDiagonalPartitioned.h:286-300): A_pp x_p."""
        x = to_device(x_p)
        out = _empty((self.nnp,))
        _call("gf_diagpart_dot_part", self._dph, _i32(1), _dp(x), _dp(out), _stream())
        return out

    def Solve_u(self, b_u, x_p):
        """Solve_u (MatrixDiagonalPartitioned.h:356-372): A_uu^-1 b_u."""
        b = to_device(b_u)
        out = _empty((self.nnu,))
        _call("gf_diagpart_solve_u", self._dph, _dp(b), _dp(out), _stream())
        return out

    def Solve(self, b, x=None):
        """Solve(b) / solve(b, x) (MatrixDiagonal.h:53-105, MatrixDiagonalPartitioned.h:306-337):
        the unknown entries A_uu^-1 b_u written over a copy of x (zeros when x is not given; the
        reference leaves the prescribed entries of Solve(b) uninitialised)."""
        torch = _torch()
        b = to_device(b)
        out = torch.zeros_like(b) if x is None else to_device(x).clone()
        _call("gf_diagpart_apply_sel", self._dph, _i32(0), _i32(0 if b.dim() == 1 else 1), _dp(b), _dp(out),
              _stream())
        return out

    def Reaction(self, x, b):
        """Reaction(x, b) (Matrix.h:490-514, MatrixDiagonalPartitioned.h:378-403): copy of b with
        its prescribed entries A_pp x_p."""
        x = to_device(x)
        out = to_device(b).clone()
        _call("gf_diagpart_apply_sel", self._dph, _i32(1), _i32(0 if x.dim() == 1 else 1), _dp(x), _dp(out),
              _stream())
        return out

    def Reaction_p(self, x_u, x_p):
        """Reaction_p (MatrixDiagonalPartitioned.h:405-420): A_pp x_p."""
        return self.Dot_p(x_u, x_p)


class Matrix(_Handle):
    """GooseFEM::Matrix (Matrix.h:653-860) as a device CSR."""

    _destroy = "gf_matrix_destroy"

    def __init__(self, conn, dofs):
        self._vec = conn if isinstance(conn, Vector) else Vector(conn, dofs)
        h = _vp()
        _call("gf_matrix_create", ctypes.byref(h), self._vec._h)
        self._h = h
        nr, nz = _i64(), _i64()
        _call("gf_matrix_sizes", h, ctypes.byref(nr), ctypes.byref(nz))
        self.nrows, self.nnz = nr.value, nz.value

    def assemble(self, elemmat):
        """Matrix::assemble; FP32 element matrices are widened entry by entry and summed in FP64."""
        f32 = _is_f32(elemmat)
        em = to_device(elemmat, _torch().float32 if f32 else None)
        _call("gf_matrix_assemble_f32" if f32 else "gf_matrix_assemble", self._h, _dp(em), _stream())

    def Dot(self, x, out=None):
        x = to_device(x)
        if out is None:
            out = _empty(tuple(x.shape))
        fn = "gf_matrix_dot_nodevec" if x.dim() == 2 else "gf_matrix_dot_dofval"
        _call(fn, self._h, _dp(x), _dp(out), _stream())
        return out

    def Todense(self):
        out = _empty((self.nrows, self.nrows))
        _call("gf_matrix_todense", self._h, _dp(out), _stream())
        return out

    def csr(self):
        """(rowptr, col, val) as GPU tensors (copies)."""
        torch = _torch()
        rp, cp, vp = _vp(), _vp(), _vp()
        _call("gf_matrix_csr", self._h, ctypes.byref(rp), ctypes.byref(cp), ctypes.byref(vp))
        rowptr = torch.empty(self.nrows + 1, dtype=torch.int32, device=device())
        col = torch.empty(self.nnz, dtype=torch.int32, device=device())
        val = _empty((self.nnz,))
        _hip_memcpy_d2d(rowptr, rp, 4 * (self.nrows + 1))
        _hip_memcpy_d2d(col, cp, 4 * self.nnz)
        _hip_memcpy_d2d(val, vp, 8 * self.nnz)
        return rowptr, col, val


class MatrixPartitioned(_Handle):
    """GooseFEM::MatrixPartitioned (MatrixPartitioned.h:34-523): the blocks A_uu, A_up, A_pu, A_pp as
    device CSR in the partitioned numbering.  assemble is bit-exact with the reference's triplet
    sums; Dot / Reaction sum each row over the first block's columns, then the second's."""

    _destroy = "gf_matpart_destroy"

    def __init__(self, conn, dofs=None, iip=None):
        self._vec = conn if isinstance(conn, VectorPartitioned) else VectorPartitioned(conn, dofs, iip)
        h = _vp()
        _call("gf_matpart_create", ctypes.byref(h), self._vec._ph)
        self._h = h
        nnz = (ctypes.c_int64 * 4)()
        _call("gf_matpart_sizes", h, nnz)
        self.nnz_blocks = tuple(nnz)
        self.nnu, self.nnp = self._vec.nnu, self._vec.nnp
        self.ndof = self._vec.ndof

    @property
    def iiu(self):
        return self._vec.iiu

    @property
    def iip(self):
        return self._vec.iip

    def assemble(self, elemmat):
        """assemble(elemmat) (MatrixPartitioned.h:156-210); FP32 element matrices are widened."""
        f32 = _is_f32(elemmat)
        em = to_device(elemmat, _torch().float32 if f32 else None)
        _call("gf_matpart_assemble_f32" if f32 else "gf_matpart_assemble", self._h, _dp(em), _stream())

    def csr(self, block):
        """(rowptr, col, val) of block 0 (uu), 1 (up), 2 (pu) or 3 (pp) as GPU tensors (copies)."""
        torch = _torch()
        rp, cp, vp = _vp(), _vp(), _vp()
        _call("gf_matpart_csr", self._h, _i32(block), ctypes.byref(rp), ctypes.byref(cp), ctypes.byref(vp))
        nrows = self.nnu if block < 2 else self.nnp
        nnz = self.nnz_blocks[block]
        rowptr = torch.empty(nrows + 1, dtype=torch.int32, device=device())
        col = torch.empty(nnz, dtype=torch.int32, device=device())
        val = _empty((nnz,))
        _hip_memcpy_d2d(rowptr, rp, 4 * (nrows + 1))
        if nnz:
            _hip_memcpy_d2d(col, cp, 4 * nnz)
            _hip_memcpy_d2d(val, vp, 8 * nnz)
        return rowptr, col, val

    def _sparse(self, block):
        torch = _torch()
        rowptr, col, val = self.csr(block)
        ncols = self.nnu if block in (0, 2) else self.nnp
        return torch.sparse_csr_tensor(rowptr, col, val, (rowptr.numel() - 1, ncols))

    @property
    def data_uu(self):
        """data_uu() (MatrixPartitioned.h:126-129) as a torch sparse CSR tensor (copy)."""
        return self._sparse(0)

    @property
    def data_up(self):
        return self._sparse(1)

    @property
    def data_pu(self):
        return self._sparse(2)

    @property
    def data_pp(self):
        return self._sparse(3)

    def Todense(self):
        """Todense() (MatrixPartitioned.h:324-350): the unpartitioned dense matrix [ndof, ndof]."""
        full = _vp()
        _call("gf_matpart_matrix", self._h, ctypes.byref(full))
        out = _empty((self.ndof, self.ndof))
        _call("gf_matrix_todense", full, _dp(out), _stream())
        return out

    def Dot(self, x, out=None):
        """Dot(x) (MatrixPartitioned.h:354-396) of a dofval or nodevec."""
        x = to_device(x)
        if out is None:
            out = _empty(tuple(x.shape))
        _call("gf_matpart_apply", self._h, _i32(0), _i32(0 if x.dim() == 1 else 1), _dp(x), _dp(out), _stream())
        return out

    def Dot_u(self, x_u, x_p):
        """B_u = A_uu x_u + A_up x_p."""
        xu, xp = to_device(x_u), to_device(x_p)
        out = _empty((self.nnu,))
        _call("gf_matpart_dot_part", self._h, _i32(0), _dp(xu), _dp(xp), _dp(out), _stream())
        return out

    def Dot_p(self, x_u, x_p):
        """B_p = A_pu x_u + A_pp x_p."""
        xu, xp = to_device(x_u), to_device(x_p)
        out = _empty((self.nnp,))
        _call("gf_matpart_dot_part", self._h, _i32(1), _dp(xu), _dp(xp), _dp(out), _stream())
        return out

    def Reaction(self, x, b):
        """Reaction(x, b) (Matrix.h:490-514, MatrixPartitioned.h:401-433): copy of b with its
        prescribed entries A_pu x_u + A_pp x_p."""
        x = to_device(x)
        out = to_device(b).clone()
        _call("gf_matpart_apply", self._h, _i32(1), _i32(0 if x.dim() == 1 else 1), _dp(x), _dp(out), _stream())
        return out

    def Reaction_p(self, x_u, x_p):
        """Reaction_p(x_u, x_p) (Matrix.h:553-575, MatrixPartitioned.h:436-447)."""
        return self.Dot_p(x_u, x_p)


def _ncols(M):
    """Column count of a dense array or a (rowptr, col, val, ncols) CSR tuple."""
    if isinstance(M, tuple):
        assert len(M) == 4, "CSR tying matrices are passed as (rowptr, col, val, ncols)"
        return int(M[3])
    return int(np.shape(M)[1])


def _csr_of(M, ncol=None):
    """Host CSR (rowptr int64, col int64, val float64) of a dense array or a (rowptr, col, val[,
    ncols]) tuple."""
    if isinstance(M, tuple):
        rp, cl, v = M[:3]
        return _i64a(rp), _i64a(cl), np.ascontiguousarray(v, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    rows, cols = np.nonzero(M)
    rp = np.zeros(M.shape[0] + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp), cols.astype(np.int64), np.ascontiguousarray(M[rows, cols])


def _csr_transpose(rp, cl, v, ncol):
    order = np.lexsort((np.repeat(np.arange(rp.size - 1), np.diff(rp)), cl))
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))[order]
    trp = np.zeros(ncol + 1, dtype=np.int64)
    np.add.at(trp, cl[order] + 1, 1)
    return np.cumsum(trp), rows.astype(np.int64), v[order]


class VectorPartitionedTyings(Vector):
    """GooseFEM::VectorPartitionedTyings (VectorPartitionedTyings.h:26-251): a Vector over the tied
    numbering [u | p | d] of Tyings::Periodic::dofs(), with the dependent DOFs d tied to the
    independent ones i = [u | p] by C_di = [C_du, C_dp].  The tying matrices may be dense arrays (as
    Tyings.Periodic.Cdu ...) or host CSR triples."""

    def __init__(self, conn, dofs, Cdu, Cdp, Cdi):
        super().__init__(conn, dofs)
        Cdu, Cdp, Cdi = (np.asarray(C, dtype=np.float64) if not isinstance(C, tuple) else C for C in (Cdu, Cdp, Cdi))
        self._du = _csr_of(Cdu)
        self._dp = _csr_of(Cdp)
        di = _csr_of(Cdi)
        self.nnd = self._du[0].size - 1
        self.nnu = _ncols(Cdu)
        self.nnp = _ncols(Cdp)
        self.nni = self.nnu + self.nnp
        assert self.nni + self.nnd == self.ndof, "tying matrices do not match the DOF numbering"
        torch = _torch()
        trp, tcl, tv = _csr_transpose(*di, self.nni)  # C_id, rows i
        self._id = (torch.as_tensor(trp.astype(np.int32), device=device()),
                    torch.as_tensor(tcl.astype(np.int32), device=device()),
                    torch.as_tensor(tv, device=device()))

    @property
    def iiu(self):
        return np.arange(self.nnu)

    @property
    def iip(self):
        return np.arange(self.nnu, self.nni)

    @property
    def iii(self):
        return np.arange(self.nni)

    @property
    def iid(self):
        return np.arange(self.nni, self.ndof)

    def AsDofs_i(self, nodevec, apply_tyings=True):
        """AsDofs_i (VectorPartitionedTyings.h:209-251): the independent DOFs of a nodevec, plus
        C_id times its dependent DOFs (the tied contributions) unless apply_tyings is False."""
        d = self.AsDofs(nodevec)
        out = d[: self.nni].clone()
        if apply_tyings and self.nnd:
            rp, cl, v = self._id
            xd = d[self.nni:].contiguous()
            _call("gf_csr_addmv", _i64(self.nni), _vp(rp.data_ptr()), _vp(cl.data_ptr()), _vp(v.data_ptr()), _dp(xd),
                  _dp(out), _stream())
        return out

    def asDofs_i(self, nodevec, dofval_i, apply_tyings=True):
        return _inplace(dofval_i, lambda r: r.copy_(self.AsDofs_i(nodevec, apply_tyings)))

    def copy_p(self, dofval_src, dofval_dest):
        """copy_p (VectorPartitionedTyings.h:188-199): dofval_dest[iip] = dofval_src[iip]."""
        src = dofval_src
        if isinstance(dofval_dest, np.ndarray):
            dofval_dest[self.nnu:self.nni] = np.asarray(_np_of(src))[self.nnu:self.nni]
        else:
            dofval_dest[self.nnu:self.nni] = to_device(src)[self.nnu:self.nni]
        return dofval_dest


class MatrixPartitionedTyings(_Handle):
    """GooseFEM::MatrixPartitionedTyings (MatrixPartitionedTyings.h:34-523) for any mesh and element
    matrices: K assembled as one CSR over the tied numbering, the condensed operator applied
    matrix-free inside MatrixPartitionedTyingsSolver's CG."""

    _destroy = "gf_mattyings_destroy"

    def __init__(self, conn, dofs, Cdu, Cdp):
        self._vec = Vector(conn, dofs)
        Cdu = np.asarray(Cdu, dtype=np.float64) if not isinstance(Cdu, tuple) else Cdu
        Cdp = np.asarray(Cdp, dtype=np.float64) if not isinstance(Cdp, tuple) else Cdp
        self._du, self._dp = _csr_of(Cdu), _csr_of(Cdp)
        self.nnd = self._du[0].size - 1
        ndof = self._vec.ndof
        self.nnp = _ncols(Cdp)
        self.nnu = _ncols(Cdu)
        assert self.nnu + self.nnp + self.nnd == ndof, "tying matrices do not match the DOF numbering"
        self.nni = self.nnu + self.nnp
        self.ndof = ndof
        h = _vp()
        _call("gf_mattyings_create", ctypes.byref(h), self._vec._h, _i64(self.nnu), _i64(self.nnp), _i64(self.nnd),
              _hp(self._du[0]), _hp(self._du[1]), _hp(self._du[2]), _hp(self._dp[0]), _hp(self._dp[1]),
              _hp(self._dp[2]))
        self._h = h

    def assemble(self, elemmat):
        """assemble(elemmat) (MatrixPartitionedTyings.h:210-305)."""
        em = to_device(elemmat)
        _call("gf_mattyings_assemble", self._h, _dp(em), _stream())

    def Todense(self):
        K = _vp()
        _call("gf_mattyings_matrix", self._h, ctypes.byref(K))
        out = _empty((self.ndof, self.ndof))
        _call("gf_matrix_todense", K, _dp(out), _stream())
        return out


class MatrixPartitionedTyingsSolver:
    """GooseFEM::MatrixPartitionedTyingsSolver (MatrixPartitionedTyings.h:527-675): x_u from the
    condensed system by an FP64 Jacobi-PCG (tol relative), x_d from the tyings; K' must be SPD."""

    def __init__(self, tol=1e-12, maxiter=None):
        self.tol = tol
        self.maxiter = maxiter
        self.iterations = 0
        self.relres = 0.0

    def solve(self, A, b, x):
        """solve(A, b, x) (:540-564) on nodevecs: x's u and d entries solved in place."""
        bd = to_device(b)
        it, res = _i64(), _f64()
        maxiter = self.maxiter or max(100, 10 * A.nnu)

        def run(r):
            _call("gf_mattyings_solve", A._h, _dp(bd), _dp(r), _f64(self.tol), _i64(maxiter), ctypes.byref(it),
                  ctypes.byref(res), _stream())

        out = _inplace(x, run)
        self.iterations, self.relres = it.value, res.value
        return out

    def Solve(self, A, b, x):
        xc = to_device(x).clone()
        return self.solve(A, b, xc)


class MatrixPartitionedSolver:
    """GooseFEM::MatrixPartitionedSolver (MatrixPartitioned.h:526-640) for a symmetric positive
    definite A_uu: x_u = A_uu^-1 (b_u - A_up x_p) by an FP64 Jacobi-preconditioned CG on the GPU
    (relative residual tol) instead of a sparse LDLT factorisation."""

    def __init__(self, tol=1e-12, maxiter=None):
        self.tol = tol
        self.maxiter = maxiter
        self.iterations = 0
        self.relres = 0.0

    def _run(self, fn, A, *args):
        it, res = _i64(), _f64()
        maxiter = self.maxiter or max(100, 10 * A.nnu)
        _call(fn, A._h, *args, _f64(self.tol), _i64(maxiter), ctypes.byref(it), ctypes.byref(res), _stream())
        self.iterations, self.relres = it.value, res.value

    def Solve(self, A, b, x, warm=False):
        """Solve(A, b, x) (MatrixPartitioned.h:583-613): copy of x with its unknown entries solved."""
        b = to_device(b)
        out = to_device(x).clone()
        self._run("gf_matpart_solve", A, _i32(0 if b.dim() == 1 else 1), _dp(b), _dp(out), _i32(int(warm)))
        return out

    def solve(self, A, b, x, warm=False):
        """solve(A, b, x): x's unknown entries solved in place (FP64 GPU tensor or numpy array)."""
        b = to_device(b)
        kind = _i32(0 if b.dim() == 1 else 1)
        return _inplace(x, lambda r: self._run("gf_matpart_solve", A, kind, _dp(b), _dp(r), _i32(int(warm))))

    def Solve_u(self, A, b_u, x_p, x_u0=None):
        """Solve_u(A, b_u, x_p) (MatrixPartitioned.h:545-560, :615-631); x_u0: optional initial guess."""
        bu, xp = to_device(b_u), to_device(x_p)
        out = _torch().zeros(A.nnu, dtype=_torch().float64, device=device()) if x_u0 is None else to_device(x_u0).clone()
        self._run("gf_matpart_solve_u", A, _dp(bu), _dp(xp), _dp(out), _i32(0 if x_u0 is None else 1))
        return out


# -------------------------------------------------------------------------------------------------
# Material laws per quadrature point, with the API the examples use from GMat (Eps, Sig, C,
# refresh, increment)
# -------------------------------------------------------------------------------------------------


class _MaterialArrays:
    def _init_arrays(self, shape):
        torch = _torch()
        self.shape = tuple(shape)
        self._Eps = torch.zeros(self.shape + (3, 3), dtype=torch.float64, device=device())
        self._Sig = torch.zeros_like(self._Eps)
        self._C = torch.zeros(self.shape + (3, 3, 3, 3), dtype=torch.float64, device=device())

    @property
    def Eps(self):
        """Strain [..., 3, 3] (device, written in place by e.g. quad.symGradN_vector(ue, mat.Eps))."""
        return self._Eps

    @Eps.setter
    def Eps(self, value):
        self._Eps.copy_(to_device(value).reshape(self._Eps.shape))

    @property
    def Sig(self):
        return self._Sig

    @property
    def C(self):
        return self._C


class Elastic(_MaterialArrays):
    """Linear elastic law of the examples (GMatElastic.Cartesian3d.Elastic2d(K, G),
    docs/examples/statics/partial-periodic_elastic.py:63): Sig = K tr(Eps) I + 2G dev(Eps)."""

    def __init__(self, K, G):
        self.K = to_device(K)
        self.G = to_device(_torch().as_tensor(np.broadcast_to(_np_of(G), tuple(self.K.shape)).copy()))
        self._init_arrays(self.K.shape)
        _call("gf_material_elastic_tangent", _i64(self.K.numel()), _i32(3), _dp(self.K), _dp(self.G), _dp(self._C),
              _stream())

    def refresh(self, compute_tangent=True):
        _call("gf_material_elastic_stress", _i64(self.K.numel()), _i32(3), _dp(self._Eps), _dp(self.K), _dp(self.G),
              _dp(self._Sig), _i32(_dtype_code(self._Eps)), _stream())

    def increment(self):
        pass


class LinearHardening(_MaterialArrays, _Handle):
    """J2 elasto-plasticity with linear isotropic hardening on the GPU, with the API of
    GMatElastoPlastic.Cartesian3d.LinearHardening2d(K, G, sigy0, H) as
    docs/examples/statics/periodic_elastoplastic.py:71-104 drives it: set Eps, refresh(), read Sig
    and C; increment() commits the plastic state (Epsp, epsp)."""

    _destroy = "gf_plastic_destroy"

    def __init__(self, K, G, sigy0, H):
        torch = _torch()
        self.K = to_device(K)
        shape = tuple(self.K.shape)

        def full(a):
            return to_device(torch.as_tensor(np.broadcast_to(_np_of(a), shape).copy()))

        self.G, self.sigy0, self.H = full(G), full(sigy0), full(H)
        self._init_arrays(shape)
        h = _vp()
        _call("gf_plastic_create", ctypes.byref(h), _i64(self.K.numel()), _dp(self.K), _dp(self.G), _dp(self.sigy0),
              _dp(self.H), _stream())
        self._h = h

    def refresh(self, compute_tangent=True):
        _call("gf_plastic_refresh", self._h, _dp(self._Eps), _dp(self._Sig), _dp(self._C) if compute_tangent else None,
              _stream())

    def increment(self):
        _call("gf_plastic_increment", self._h, _stream())

    def newton_q4(self, quad, elemvec, fe=None, Ke=None, Sig=None):
        """Fused pass of the statics Newton iteration on a Quad4 QuadraturePlanar (4 points):
        strain from elemvec, refresh, element force and stiffness, without storing Eps or C.
        Equal bit for bit to symGradN_vector -> refresh() -> Int_gradN_dot_tensor2_dV(Sig),
        Int_gradN_dot_tensor4_dot_gradNT_dV(C).  Returns (fe, Ke); Sig is written if given (and
        then also left in self.Sig when it is that tensor)."""
        ev = to_device(elemvec)
        ne = quad.nelem
        fe = _out(fe, (ne, 4, 2))
        Ke = _out(Ke, (ne, 8, 8))
        _call("gf_plastic_newton_q4", self._h, quad._h, _dp(ev), _dp(fe), _dp(Ke), _dp(Sig), _stream())
        return fe, Ke

    def _state(self, which, shape):
        ptrs = [_vp(), _vp(), _vp(), _vp()]
        _call("gf_plastic_state", self._h, *[ctypes.byref(p) for p in ptrs])
        out = _empty(shape)
        _hip_memcpy_d2d(out, ptrs[which], 8 * out.numel())
        return out

    @property
    def Epsp(self):
        """Plastic strain [..., 3, 3] (copy)."""
        return self._state(0, self.shape + (3, 3))

    @property
    def epsp(self):
        """Equivalent plastic strain [...] (copy)."""
        return self._state(1, self.shape)


Elastic2d = Elastic
LinearHardening2d = LinearHardening


# -------------------------------------------------------------------------------------------------
# Explicit dynamics (fused velocity-Verlet step)
# -------------------------------------------------------------------------------------------------


class ExplicitQuad4(_Handle):
    """Fused velocity-Verlet integrator of docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp
    on a Quad4 plane-strain mesh with per-element moduli (K_e, G_e as float32)."""

    _destroy = "gf_explicit_destroy"

    def __init__(self, vector: Vector, coor, K_e, G_e, inv_mass, dt, fext=None):
        torch = _torch()
        self._vec = vector
        self._coor = to_device(coor)
        self._K = to_device(K_e, dtype=torch.float32)
        self._G = to_device(G_e, dtype=torch.float32)
        self._inv = to_device(inv_mass)
        self._fext = None if fext is None else to_device(fext)
        h = _vp()
        _call("gf_explicit_create", ctypes.byref(h), vector._h, _dp(self._coor), _dp(self._K), _dp(self._G),
              _dp(self._inv), _dp(self._fext), _f64(dt))
        self._h = h

    @property
    def structured(self):
        s = _i32()
        _call("gf_explicit_is_structured", self._h, ctypes.byref(s))
        return bool(s.value)

    @property
    def path(self):
        """'structured' (tile kernel), 'blocks' (general mesh, fused per node block) or 'gather'
        (general mesh with tied DOFs: element kernel + ordered per-DOF gather)."""
        s = _i32()
        _call("gf_explicit_path", self._h, ctypes.byref(s))
        return ("structured", "blocks", "gather")[s.value]

    def _state_ptrs(self):
        u, v, a = _vp(), _vp(), _vp()
        _call("gf_explicit_state", self._h, ctypes.byref(u), ctypes.byref(v), ctypes.byref(a))
        return u, v, a

    def set_state(self, u, v, a):
        n = self._vec.ndof
        for ptr, src in zip(self._state_ptrs(), (u, v, a)):
            src = to_device(src)
            assert src.numel() == n
            hip = _hip_runtime()
            rc = hip.hipMemcpyAsync(ptr, _vp(src.data_ptr()), ctypes.c_size_t(8 * n), 3, _stream())
            if rc:
                raise GooseFEMError(GF_EHIP, "hipMemcpyAsync")
        _torch().cuda.current_stream().synchronize()

    def state(self):
        n = self._vec.ndof
        outs = []
        for ptr in self._state_ptrs():
            o = _empty((n,))
            _hip_memcpy_d2d(o, ptr, 8 * n)
            outs.append(o)
        return tuple(outs)

    def step(self, nsteps=1, precision=F64, store_fint=False, stream=None):
        s = _stream() if stream is None else stream
        _call("gf_explicit_step", self._h, _i64(nsteps), _i32(precision), _i32(int(store_fint)), s)

    def fint(self):
        p = _vp()
        _call("gf_explicit_fint", self._h, ctypes.byref(p))
        o = _empty((self._vec.ndof,))
        _hip_memcpy_d2d(o, p, 8 * self._vec.ndof)
        return o

    def energy(self, mass, out=None):
        """(Ekin, Epot) of the current state as a device tensor of 2 (written into out[0:2] if given):
        Ekin = sum_d mass_d v_d^2 / 2 (mass: the lumped mass of every DOF), Epot = u . f_int / 2 of
        the last step run with store_fint (Elastic-VelocityVerlet/main.cpp:155-160)."""
        if not hasattr(self, "_mass") or self._mass_src is not mass:
            self._mass, self._mass_src = to_device(mass), mass
        out = _empty((2,)) if out is None else out
        _call("gf_explicit_energy", self._h, _dp(self._mass), _dp(out), _vp(out.data_ptr() + 8), _stream())
        return out


# -------------------------------------------------------------------------------------------------
# Tied-system PCG (MatrixPartitionedTyings operator, FP64 Jacobi-PCG)
# -------------------------------------------------------------------------------------------------


class TyingsPCG(_Handle):
    """FP64 Jacobi-PCG on K' = A_uu + A_ud C_du + C_du^T A_du + C_du^T A_dd C_du
    (MatrixPartitionedTyings.h:655-674), matrix-free for a linear elastic material.

    vector: Vector(conn, tyings.dofs); quad: QuadraturePlanar / Hex8 Quadrature (Gauss) of the mesh;
    K_e, G_e: per-element moduli; tyings: Tyings.Periodic.
    """

    _destroy = "gf_pcg_destroy"

    def __init__(self, vector, quad, K_e, G_e, tyings):
        self._vec, self._quad = vector, quad
        self._K = to_device(K_e)
        self._G = to_device(G_e)
        self.nnu, self.nnp = tyings.nnu, tyings.nnp
        self.nnd = tyings.nnd
        self._cdu = tuple(np.ascontiguousarray(a) for a in tyings.Cdu_csr)
        self._cdp = tuple(np.ascontiguousarray(a) for a in tyings.Cdp_csr)
        h = _vp()
        _call("gf_pcg_create", ctypes.byref(h), vector._h, quad._h, _dp(self._K), _dp(self._G), _i64(self.nnu),
              _i64(self.nnp), _hp(self._cdu[0]), _hp(self._cdu[1]), _hp(self._cdu[2]), _hp(self._cdp[0]),
              _hp(self._cdp[1]), _hp(self._cdp[2]))
        self._h = h

    @property
    def structured(self):
        """True when K' x_u runs in the fused structured periodic kernel."""
        f = _i32(0)
        _call("gf_pcg_is_structured", self._h, ctypes.byref(f))
        return bool(f.value)

    def set_moduli(self, K_e, G_e):
        """Replace the per-element moduli (same mesh and tyings) and refresh the Jacobi diagonal."""
        self._K = to_device(K_e)
        self._G = to_device(G_e)
        _call("gf_pcg_set_moduli", self._h, _dp(self._K), _dp(self._G), _stream())

    def apply(self, x_u=None, x_p=None):
        """K' x_u + K'_up x_p."""
        xu = None if x_u is None else to_device(x_u)
        xp = None if x_p is None else to_device(x_p)
        out = _empty((self.nnu,))
        _call("gf_pcg_apply", self._h, _dp(xu), _dp(xp), _dp(out), _stream())
        return out

    def expand(self, x_u=None, x_p=None):
        """Full dofval [x_u, x_p, C_du x_u + C_dp x_p]."""
        xu = None if x_u is None else to_device(x_u)
        xp = None if x_p is None else to_device(x_p)
        out = _empty((self.nnu + self.nnp + self.nnd,))
        _call("gf_pcg_expand", self._h, _dp(xu), _dp(xp), _dp(out), _stream())
        return out

    @property
    def inverse_diagonal(self):
        p = _vp()
        _call("gf_pcg_inverse_diagonal", self._h, ctypes.byref(p))
        out = _empty((self.nnu,))
        _hip_memcpy_d2d(out, p, 8 * self.nnu)
        return out

    def solve(self, b_u, x0=None, tol=1e-10, maxiter=100000, check_every=32, raise_on_fail=True):
        """Returns (x_u, iterations, true relative residual)."""
        b = to_device(b_u)
        x = _empty((self.nnu,)) if x0 is None else to_device(x0).clone()
        if x0 is None:
            x.zero_()
        it, res = _i64(), _f64()
        rc = _fn("gf_pcg_solve")(self._h, _dp(b), _dp(x), _f64(tol), _i64(maxiter), _i64(check_every), ctypes.byref(it),
               ctypes.byref(res), _stream())
        if rc != GF_OK and (raise_on_fail or rc != GF_ENOTCONV):
            _check(rc)
        return x, it.value, res.value

    @property
    def start_used(self):
        """Start of the last solve: GF_START_GIVEN (the caller's x0), GF_START_PREVIOUS (the
        previous converged solution, taken when x0's residual was not finite or above ||b||) or
        GF_START_ZERO."""
        v = _i32(0)
        _call("gf_pcg_start_used", self._h, ctypes.byref(v))
        return v.value
