"""Hex8 element stiffness on the GPU: FP64 path bit-exact with the oracle's reference-order loop,
BF16 hi/lo MFMA path within a stated tolerance, and the assembled CSR product equal to the
matrix-free sum_e K_e u_e."""

import numpy as np
import pytest

import oracle
from oracle import HEX8

pytestmark = pytest.mark.gpu

# BF16x3 (hi*hi + hi*lo + lo*hi, FP32 accumulation) vs FP64: norm-wise relative error bound
TOL_BF16X3 = 1e-5


@pytest.fixture(scope="module")
def gf():
    import goosefem_amd

    goosefem_amd.lib()
    return goosefem_amd


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def hex8_problem(n=4, distort=0.0, seed=0):
    c, conn = oracle.hex8_regular(n, n, n)
    if distort:
        c = c + distort * np.random.default_rng(seed).uniform(-1, 1, c.shape)
    vec = oracle.Vector(conn, oracle.dofs(c.shape[0], 3))
    x = vec.AsElement(c)
    q = oracle.Quadrature(HEX8, x)
    rng = np.random.default_rng(seed + 1)
    K = np.full((q.nelem, 8), 10.0)
    G = rng.uniform(0.5, 1.5, (q.nelem, 8))
    C = oracle.elastic_tangent(K, G)
    return c, conn, x, q, C


@pytest.mark.parametrize("distort", [0.0, 0.15])
def test_stiffness_fp64_and_bf16x3(gf, distort):
    c, conn, x, q, C = hex8_problem(5, distort)
    ref = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    gq = gf.Element.Hex8.Quadrature(x)
    k64 = host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C, gf.F64))
    assert np.array_equal(k64, ref)
    kbf = host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C, gf.BF16X3))
    err = relerr(kbf, ref)
    assert err < TOL_BF16X3, err
    # symmetric and rigid-body free to the same tolerance
    assert relerr(np.swapaxes(kbf, 1, 2), ref) < TOL_BF16X3
    rigid = np.tile(np.array([1.0, 0.0, 0.0]), (q.nelem, 8, 1)).reshape(q.nelem, 24)
    r = np.einsum("eab,eb->ea", kbf, rigid)
    assert np.abs(r).max() < TOL_BF16X3 * np.abs(ref).max() * 24


def test_stiffness_bf16x3_general_tensor(gf):
    # no symmetry assumed: random C_ijkl
    c, conn, x, q, _ = hex8_problem(3, 0.1, seed=5)
    C = np.random.default_rng(9).standard_normal((q.nelem, 8, 3, 3, 3, 3))
    ref = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    gq = gf.Element.Hex8.Quadrature(x)
    kbf = host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C, gf.BF16X3))
    assert relerr(kbf, ref) < TOL_BF16X3


@pytest.mark.parametrize("periodic", [False, True])
def test_assembled_dot_equals_matrix_free(gf, periodic):
    n = 4
    c, conn, x, q, C = hex8_problem(n)
    dofs = (oracle.dofs_periodic(c.shape[0], 3, oracle.hex8_nodes_periodic(n, n, n)) if periodic
            else oracle.dofs(c.shape[0], 3))
    gq = gf.Element.Hex8.Quadrature(x)
    Ke = gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C, gf.F64)
    vec = gf.Vector(conn, dofs)
    A = gf.Matrix(vec, None)
    A.assemble(Ke)
    u = np.random.default_rng(3).standard_normal(vec.ndof)
    Ku = host(A.Dot(u))
    mf = host(vec.AssembleDofs(gf.elemmat_dot(Ke, vec.AsElement(u))))
    assert relerr(Ku, mf) < 1e-14
    # against the oracle's CSR product (bit-exact: same ordered sums)
    rowptr, col, val = oracle.matrix_assemble(conn, dofs, host(Ke))
    assert np.array_equal(Ku, oracle.csr_dot(rowptr, col, val, u))


@pytest.mark.parametrize("K,nelem", [(24, 37), (8, 101), (6, 5)])
def test_elemmat_dot_matches_serial_row_sums(gf, K, nelem):
    """Batched element product out[e] = K_e u_e: FP64, rows summed in column order (the serial loop),
    for the wave-tiled kernels (K = 8, 24; row counts not multiples of 64) and the generic one."""
    rng = np.random.default_rng(K)
    Ke = rng.standard_normal((nelem, K, K))
    ue = rng.standard_normal((nelem, K))
    got = host(gf.elemmat_dot(Ke, ue))
    ref = np.zeros((nelem, K))
    for e in range(nelem):
        for i in range(K):
            s = 0.0
            for c in range(K):
                s += Ke[e, i, c] * ue[e, c]
            ref[e, i] = s
    assert np.array_equal(got.reshape(nelem, K), ref)


@pytest.mark.parametrize("distort", [0.0, 0.15])
def test_stiffness_bf16x3_fp32_io(gf, distort):
    # FP32 tangent in, FP32 element matrices out.  The MFMA kernel rounds its FP64 input to FP32
    # before the split, so fed the same FP32 values through the FP64 interface it must give the
    # FP32 interface's result widened, bit for bit; both stay within the BF16x3 bound of FP64.
    import torch

    c, conn, x, q, _ = hex8_problem(5, distort, seed=3)
    rng = np.random.default_rng(4)
    Kq = np.full((q.nelem, 8), 10.0)
    Gq = rng.uniform(0.5, 1.5, (q.nelem, 8))
    C32 = gf.elastic_tangent(Kq, Gq, dtype=torch.float32)
    assert C32.dtype == torch.float32
    C64 = oracle.elastic_tangent(Kq, Gq)
    assert np.array_equal(host(C32), C64.astype(np.float32))
    gq = gf.Element.Hex8.Quadrature(x)
    k32 = gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C32, gf.BF16X3)
    assert k32.dtype == torch.float32 and tuple(k32.shape) == (q.nelem, 24, 24)
    kw = host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C32.double(), gf.BF16X3))
    assert np.array_equal(host(k32).astype(np.float64), kw)
    ref = q.Int_gradN_dot_tensor4_dot_gradNT_dV(host(C32).astype(np.float64))
    assert relerr(host(k32).astype(np.float64), ref) < TOL_BF16X3


@pytest.mark.parametrize("periodic", [False, True])
def test_fp32_elemmat_assembly_and_product(gf, periodic):
    # FP32 element matrices: CSR values and the matrix-free product equal the FP64 paths fed the
    # widened values, bit for bit (widening is exact; sums keep their order)
    import torch

    n = 3
    c, conn = oracle.hex8_regular(n, n, n)
    if periodic:
        dofs = oracle.dofs_periodic(c.shape[0], 3, oracle.hex8_nodes_periodic(n, n, n))
    else:
        dofs = oracle.dofs(c.shape[0], 3)
    rng = np.random.default_rng(9)
    a32 = rng.standard_normal((conn.shape[0], 24, 24)).astype(np.float32)
    a64 = a32.astype(np.float64)
    rowptr, col, val = oracle.matrix_assemble(conn, dofs, a64)
    A = gf.Matrix(conn, dofs)
    A.assemble(torch.from_numpy(a32).cuda())
    grp, gcol, gval = (host(t) for t in A.csr())
    assert np.array_equal(grp, rowptr) and np.array_equal(gcol, col) and np.array_equal(gval, val)
    ue = rng.standard_normal((conn.shape[0], 24))
    k32 = torch.from_numpy(a32).cuda()
    assert np.array_equal(host(gf.elemmat_dot(k32, ue)), host(gf.elemmat_dot(a64, ue)))
    # the row-serial kernels (rows longer than 256 entries) on FP32 input
    nfan = 70
    fconn = np.array([[0, 1 + 3 * i, 2 + 3 * i, 3 + 3 * i] for i in range(nfan)], dtype=np.int64)
    fdofs = oracle.dofs(3 * nfan + 1, 2)
    f32 = rng.standard_normal((nfan, 8, 8)).astype(np.float32)
    rp, cl, vl = oracle.matrix_assemble(fconn, fdofs, f32.astype(np.float64))
    B = gf.Matrix(fconn, fdofs)
    B.assemble(torch.from_numpy(f32).cuda())
    assert np.array_equal(host(B.csr()[2]), vl)


def test_alignment_rules(gf):
    # the FP32-I/O stiffness needs a 16-byte aligned tensor4 (clear error otherwise); the element
    # GEMV and the FP64 stiffness fall back to their per-row / per-lane stores for unaligned arrays
    # and give the same results
    import torch

    c, conn, x, q, C = hex8_problem(2, 0.1, seed=5)
    gq = gf.Element.Hex8.Quadrature(x)
    flat = torch.from_numpy(C.astype(np.float32).ravel()).cuda()
    buf = torch.zeros(flat.numel() + 1, dtype=torch.float32, device="cuda")
    buf[1:] = flat
    shifted = buf[1:].view(C.shape)
    assert shifted.data_ptr() % 16 != 0
    with pytest.raises(gf.GooseFEMError):
        gq.Int_gradN_dot_tensor4_dot_gradNT_dV(shifted, gf.BF16X3)
    rng = np.random.default_rng(6)
    Ke = rng.standard_normal((q.nelem, 24, 24))
    ue = rng.standard_normal((q.nelem, 24))
    ref = host(gf.elemmat_dot(Ke, ue))
    kb = torch.zeros(Ke.size + 1, dtype=torch.float64, device="cuda")
    kb[1:] = torch.from_numpy(Ke.ravel()).cuda()
    assert kb[1:].data_ptr() % 16 != 0
    assert np.array_equal(host(gf.elemmat_dot(kb[1:].view(Ke.shape), ue)), ref)
    # the FP64 Hex8 stiffness streams an aligned element matrix out of LDS and stores an unaligned
    # one element by element: the same bits, and the oracle's
    k64 = host(gq.Int_gradN_dot_tensor4_dot_gradNT_dV(C))
    assert np.array_equal(k64, q.Int_gradN_dot_tensor4_dot_gradNT_dV(C))
    ob = torch.zeros(k64.size + 1, dtype=torch.float64, device="cuda")
    out = ob[1:].view(k64.shape)
    assert out.data_ptr() % 16 != 0
    gq.Int_gradN_dot_tensor4_dot_gradNT_dV(torch.from_numpy(C).cuda(), out=out)
    assert np.array_equal(host(out), k64)
