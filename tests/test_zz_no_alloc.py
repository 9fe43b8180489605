"""The allocation-free entry point gf_matrix_dot_nodevec_noalloc (no counterpart in the reference;
for callers that cannot allocate device memory) against gf_matrix_dot_nodevec (Matrix::Dot on a
nodevec, Matrix.h:804-809), bit for bit, through the C ABI: two calls in a row, the second on a
smaller matrix, on periodic dofs (several nodes per DOF: the last-writer AsDofs); and its size
limit."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gf():
    import torch

    import goosefem_amd

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    goosefem_amd.lib()
    return goosefem_amd


def _system(gf, nx, ny, seed):
    import torch

    mesh = gf.Mesh.Quad4.Regular(nx, ny)
    conn, dofs = mesh.conn, mesh.dofsPeriodic
    rng = np.random.default_rng(seed)
    Ke = rng.standard_normal((mesh.nelem, 8, 8))
    Ke = Ke + np.swapaxes(Ke, 1, 2)
    A = gf.Matrix(conn, dofs)
    A.assemble(Ke)
    x = torch.from_numpy(rng.standard_normal((mesh.nnode, 2))).cuda()
    return A, x


def _noalloc(gf, A, x):
    import torch

    out = torch.full_like(x, np.nan)
    gf._call("gf_matrix_dot_nodevec_noalloc", A._h, gf._dp(x), gf._dp(out), gf._stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_dot_nodevec_noalloc_matches_allocating_entry_point(gf):
    import torch

    for nx, ny, seed in ((60, 45, 1), (7, 5, 2)):  # the second call is smaller than the first
        A, x = _system(gf, nx, ny, seed)
        ref = A.Dot(x)
        torch.cuda.synchronize()
        got = _noalloc(gf, A, x)
        assert np.array_equal(got, ref.cpu().numpy()), (nx, ny)
    # a third call repeats the first, after the smaller one (nothing stale is read)
    A, x = _system(gf, 60, 45, 1)
    assert np.array_equal(_noalloc(gf, A, x), A.Dot(x).cpu().numpy())


def test_dot_nodevec_noalloc_rejects_oversize(gf):
    import torch

    # 2 x 725^2 DOFs > 2^20: beyond the static staging
    mesh = gf.Mesh.Quad4.Regular(724, 724)
    A = gf.Matrix(mesh.conn, mesh.dofs)
    A.assemble(np.zeros((mesh.nelem, 8, 8)))
    x = torch.zeros((mesh.nnode, 2), dtype=torch.float64, device="cuda")
    out = torch.empty_like(x)
    with pytest.raises(gf.GooseFEMError) as e:
        gf._call("gf_matrix_dot_nodevec_noalloc", A._h, gf._dp(x), gf._dp(out), gf._stream())
    assert e.value.code == gf.GF_EINVAL and "exceed" in str(e.value)
