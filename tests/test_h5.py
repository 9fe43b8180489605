"""HDF5 output (libgoosefem_amd_h5.so, the H5Easy semantics of the reference's examples): the
datasets Elastic-VelocityVerlet/main.cpp:161-176 writes, round-tripped through the HDF5 C library;
the per-increment series grown entry by entry; the dump / overwrite / read-only error paths."""

import os

import numpy as np
import pytest

import goosefem_amd as gf
from goosefem_amd import h5


def test_example_output_round_trip(tmp_path):
    mesh = gf.Mesh.Quad4.Regular(6, 4)
    conn = mesh.conn.astype(np.uint64)  # size_t in the reference
    coor = mesh.coor
    disp = np.random.default_rng(0).standard_normal(coor.shape)
    path = os.fspath(tmp_path / "example.hdf5")
    with h5.File(path, h5.File.Overwrite) as f:
        for inc in range(13):  # the per-increment series of the example
            h5.dump(f, "/global/t", inc * 0.01, inc)
            h5.dump(f, "/global/Ekin", 0.5 * inc ** 2, inc)
            h5.dump(f, "/global/Epot", -1.0 * inc, inc)
        h5.dump(f, "/mesh/conn", conn)
        h5.dump(f, "/mesh/coor", coor)
        h5.dump(f, "/mesh/disp", disp)
    assert open(path, "rb").read(8) == b"\x89HDF\r\n\x1a\n"  # an HDF5 file signature
    with h5.File(path, h5.File.ReadOnly) as f:
        assert h5.exist(f, "/mesh/conn") and h5.exist(f, "/global/Ekin") and not h5.exist(f, "/mesh/nothing")
        assert not h5.exist(f, "/nogroup/x")
        assert h5.getShape(f, "/mesh/conn") == conn.shape
        assert h5.getShape(f, "/global/t") == (13,)
        got = h5.load(f, "/mesh/conn", np.uint64)
        assert got.dtype == np.uint64 and np.array_equal(got, conn)
        assert np.array_equal(h5.load(f, "/mesh/coor"), coor)
        assert np.array_equal(h5.load(f, "/mesh/disp"), disp)
        assert np.array_equal(h5.load(f, "/global/t"), np.arange(13) * 0.01)
        assert np.array_equal(h5.load(f, "/global/Ekin"), 0.5 * np.arange(13) ** 2)
        # HDF5's numeric conversion: the uint64 connectivity read as float64
        assert np.array_equal(h5.load(f, "/mesh/conn", np.float64), conn.astype(np.float64))
        with pytest.raises(h5.H5Error):
            h5.dump(f, "/mesh/more", coor)  # read-only file


def test_dump_modes(tmp_path):
    path = os.fspath(tmp_path / "modes.hdf5")
    a = np.arange(12.0).reshape(3, 4)
    with h5.File(path, "w") as f:
        h5.dump(f, "/a/b/c", a)
        with pytest.raises(h5.H5Error):
            h5.dump(f, "/a/b/c", a)  # exists: H5Easy::DumpMode::Create
        h5.dump(f, "/a/b/c", 2 * a, overwrite=True)
        with pytest.raises(h5.H5Error):
            h5.dump(f, "/a/b/c", np.zeros(5), overwrite=True)  # another shape
        h5.dump(f, "/s", 3.5)  # a scalar (rank 0)
        h5.dump(f, "/series", 1.0, 4)  # sparse series: entries 0..3 are the fill value
    with h5.File(path, "a") as f:  # ReadWrite keeps the content and grows the series
        h5.dump(f, "/series", 2.0, 6)
        assert np.array_equal(h5.load(f, "/a/b/c"), 2 * a)
        assert h5.getShape(f, "/s") == () and float(h5.load(f, "/s")) == 3.5
        s = h5.load(f, "/series")
        assert s.shape == (7,) and s[4] == 1.0 and s[6] == 2.0
    with pytest.raises(h5.H5Error):
        h5.File(os.fspath(tmp_path / "absent.hdf5"), "r")


def test_cpp_h5easy(tmp_path):
    """include/GooseFEM/H5Easy.h: the example's output code, compiled (tests/cpp/test_h5easy.cpp)."""
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "test_h5easy")
    assert os.path.exists(exe), "build it first: __graft_entry__.build() (make -C tests/cpp)"
    r = subprocess.run([exe, os.fspath(tmp_path / "cpp.hdf5")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    with h5.File(os.fspath(tmp_path / "cpp.hdf5"), "r") as f:
        assert h5.getShape(f, "/mesh/conn") == (24, 4)
        assert h5.load(f, "/mesh/conn", np.uint64).dtype == np.uint64
