"""HDF5 field output with the semantics of the H5Easy calls of the reference's examples (HighFive
H5Easy::File / dump / load / exist / getShape; docs/examples/dynamics/Elastic-VelocityVerlet/
main.cpp:161-176), over libgoosefem_amd_h5.so (the HDF5 C library).  Host numpy arrays; device
tensors are copied to the host first.

    from goosefem_amd import h5
    with h5.File("example.hdf5", h5.File.Overwrite) as f:
        h5.dump(f, "/mesh/conn", conn)          # size_t-like arrays are stored as uint64
        h5.dump(f, "/mesh/coor", coor)
        h5.dump(f, "/global/Ekin", ekin, index)  # one entry of an extendible series
    disp = h5.load(h5.File("example.hdf5", h5.File.ReadOnly), "/mesh/disp")
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgoosefem_amd_h5.so")
F64, I64, U64 = 0, 1, 2
_lib = None


class H5Error(RuntimeError):
    """A nonzero status of libgoosefem_amd_h5.so (H5Easy throws)."""


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with `python -m goosefem_amd.build` (needs the HDF5 "
                              "C library, /opt/conda by default or $GF_HDF5_DIR)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.gf_h5_last_error.restype = ctypes.c_char_p
        vp, i64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        sigs = {
            "gf_h5_open": [vp, ctypes.c_char_p, ci],
            "gf_h5_close": [vp],
            "gf_h5_exist": [vp, ctypes.c_char_p, vp],
            "gf_h5_dump": [vp, ctypes.c_char_p, vp, ci, ci, vp, ci],
            "gf_h5_dump_at": [vp, ctypes.c_char_p, vp, ci, i64],
            "gf_h5_shape": [vp, ctypes.c_char_p, vp, vp],
            "gf_h5_load": [vp, ctypes.c_char_p, vp, ci],
        }
        for name, argtypes in sigs.items():
            f = getattr(_lib, name)
            f.argtypes = argtypes
            f.restype = ctypes.c_int
    return _lib


def _check(rc):
    if rc != 0:
        raise H5Error(lib().gf_h5_last_error().decode())


class File:
    """H5Easy::File(path, mode)."""

    Overwrite, ReadOnly, ReadWrite = 0, 1, 2

    def __init__(self, path, mode=ReadOnly):
        if isinstance(mode, str):
            mode = {"w": File.Overwrite, "r": File.ReadOnly, "a": File.ReadWrite}[mode]
        h = ctypes.c_void_p()
        _check(lib().gf_h5_open(ctypes.byref(h), os.fsencode(path), int(mode)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().gf_h5_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _host(data):
    if hasattr(data, "detach"):
        data = data.detach().cpu().numpy()
    a = np.asarray(data)
    # np.array keeps rank 0 (np.ascontiguousarray would make it 1-D)
    if a.dtype.kind == "f":
        return np.array(a, dtype=np.float64, order="C"), F64
    if a.dtype.kind == "u":
        return np.array(a, dtype=np.uint64, order="C"), U64
    if a.dtype.kind in "ib":
        return np.array(a, dtype=np.int64, order="C"), I64
    raise TypeError(f"unsupported dtype {a.dtype}")


def exist(file: File, path: str) -> bool:
    """H5Easy::exist."""
    e = ctypes.c_int(0)
    _check(lib().gf_h5_exist(file._h, path.encode(), ctypes.byref(e)))
    return bool(e.value)


def dump(file: File, path: str, data, index=None, overwrite=False):
    """H5Easy::dump(file, path, data) (overwrite: H5Easy::DumpMode::Overwrite), or with index the
    entry `index` of an extendible 1-D dataset (H5Easy::dump(file, path, value, {index}))."""
    if index is not None:
        a, dt = _host(np.asarray(data).reshape(()))
        _check(lib().gf_h5_dump_at(file._h, path.encode(), a.ctypes.data_as(ctypes.c_void_p), dt, int(index)))
        return
    a, dt = _host(data)
    shape = np.asarray(a.shape, dtype=np.int64)
    _check(lib().gf_h5_dump(file._h, path.encode(), a.ctypes.data_as(ctypes.c_void_p), dt, a.ndim,
                            shape.ctypes.data_as(ctypes.c_void_p), int(bool(overwrite))))


def getShape(file: File, path: str):
    """H5Easy::getShape."""
    r = ctypes.c_int(0)
    shape = np.zeros(8, dtype=np.int64)
    _check(lib().gf_h5_shape(file._h, path.encode(), ctypes.byref(r), shape.ctypes.data_as(ctypes.c_void_p)))
    return tuple(int(s) for s in shape[: r.value])


def load(file: File, path: str, dtype=np.float64):
    """H5Easy::load<T>: the whole dataset as a numpy array of dtype (float64, int64 or uint64)."""
    dtype = np.dtype(dtype)
    code = {np.dtype(np.float64): F64, np.dtype(np.int64): I64, np.dtype(np.uint64): U64}[dtype]
    out = np.empty(getShape(file, path), dtype=dtype)
    _check(lib().gf_h5_load(file._h, path.encode(), out.ctypes.data_as(ctypes.c_void_p), code))
    return out
