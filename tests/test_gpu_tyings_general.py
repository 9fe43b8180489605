"""MatrixPartitionedTyings / MatrixPartitionedTyingsSolver / VectorPartitionedTyings for any element
matrices (gf_mattyings_*, MatrixPartitionedTyings.h:34-675, VectorPartitionedTyings.h:26-251): the
condensed solve against a dense condensation of the oracle-assembled matrix, and the reference's
periodic elasto-plastic example (docs/examples/statics/periodic_elastoplastic.py:1-147, fewer
increments) run on the GPU against the same loop on the CPU (oracle quadrature, material and a dense
direct solve)."""

import numpy as np
import pytest

import oracle
from oracle import QUAD4_PLANAR

pytestmark = pytest.mark.gpu


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
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def periodic_setup(gf, n):
    mesh = gf.Mesh.Quad4.Regular(n, n)
    control = gf.Tyings.Control(mesh.coor, mesh.dofs)
    coor, dofs = control.coor, control.dofs
    control_dofs, control_nodes = control.controlDofs, control.controlNodes
    iip = np.concatenate((control_dofs.ravel(), dofs[mesh.nodesOrigin, :].ravel()))
    tyings = gf.Tyings.Periodic(coor, dofs, control_dofs, mesh.nodesPeriodic, iip)
    return mesh, coor, tyings, control_nodes


def dense_solve(K, Cdu, Cdp, nnu, nnp, b, x):
    """MatrixPartitionedTyingsSolver::solve (MatrixPartitionedTyings.h:540-564, 655-674) densely,
    on dofvals b, x [ndof]; returns x with its u and d entries solved."""
    ni = nnu + nnp
    u, p, d = slice(0, nnu), slice(nnu, ni), slice(ni, K.shape[0])
    Cud = Cdu.T
    ACuu = K[u, u] + K[u, d] @ Cdu + Cud @ K[d, u] + Cud @ K[d, d] @ Cdu
    ACup = K[u, p] + K[u, d] @ Cdp + Cud @ K[d, p] + Cud @ K[d, d] @ Cdp
    Bu = b[u] + Cud @ b[d]
    Xu = np.linalg.solve(ACuu, Bu - ACup @ x[p])
    out = x.copy()
    out[u] = Xu
    out[d] = Cdu @ Xu + Cdp @ x[p]
    return out


def test_condensed_solve_against_dense(gf):
    mesh, coor, ty, _ = periodic_setup(gf, 6)
    dofs = ty.dofs
    conn = mesh.conn
    x = oracle.Vector(conn, dofs).AsElement(coor)
    oq = oracle.Quadrature(QUAD4_PLANAR, x)
    rng = np.random.default_rng(0)
    G = np.exp(0.4 * rng.standard_normal((oq.nelem, oq.nip)))
    Ke = oq.Int_gradN_dot_tensor4_dot_gradNT_dV(oracle.elastic_tangent(np.full_like(G, 10.0), G))
    A = gf.MatrixPartitionedTyings(conn, dofs, ty.Cdu, ty.Cdp)
    A.assemble(Ke)
    rp, cl, val = oracle.matrix_assemble(conn, dofs, Ke)
    Kd = np.zeros((rp.size - 1,) * 2)
    for r in range(rp.size - 1):
        Kd[r, cl[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    assert np.array_equal(host(A.Todense()), Kd)
    ov = oracle.Vector(conn, dofs)
    b = rng.standard_normal(coor.shape)
    xin = np.zeros_like(coor)
    xin[dofs >= ty.nnu] = 0.0
    pn = (dofs >= ty.nnu) & (dofs < ty.nni)
    xin[pn] = 1e-3 * rng.standard_normal(pn.sum())
    solver = gf.MatrixPartitionedTyingsSolver(tol=1e-12)
    got = host(solver.solve(A, b, xin.copy()))
    ref = ov.AsNode(dense_solve(Kd, ty.Cdu, ty.Cdp, ty.nnu, ty.nnp, ov.AsDofs(b), ov.AsDofs(xin)))
    assert solver.relres < 1e-11 and solver.iterations > 0
    assert np.max(np.abs(got - ref)) < 1e-8 * np.max(np.abs(ref))
    # a non-finite right-hand side fails at once (GF_EINVAL), not after the iteration limit
    bad = b.copy()
    bad[1, 0] = np.inf
    with pytest.raises(gf.GooseFEMError) as e:
        solver.solve(A, bad, xin.copy())
    assert e.value.code == gf.GF_EINVAL
    # AsDofs_i: independent DOFs plus the tied contributions C_id b_d
    vec = gf.VectorPartitionedTyings(conn, dofs, ty.Cdu, ty.Cdp, ty.Cdi)
    bd = ov.AsDofs(b)
    ref_i = bd[: ty.nni] + ty.Cdi.T @ bd[ty.nni:]
    assert np.allclose(host(vec.AsDofs_i(b)), ref_i, rtol=1e-14, atol=1e-14)
    assert np.array_equal(host(vec.AsDofs_i(b, apply_tyings=False)), bd[: ty.nni])


def run_example(gf, nsteps, on_gpu, dgamma=0.01):
    """periodic_elastoplastic.py:1-147 with nsteps increments of dgamma shear (the example takes 100
    of 0.001; larger steps reach the plastic range in a few)."""
    mesh, coor, ty, control_nodes = periodic_setup(gf, 5)
    conn, dofs = mesh.conn, ty.dofs
    nelem = mesh.nelem
    kappa, mu = np.ones((nelem, 4)), np.ones((nelem, 4))
    sigy0, H = 0.05 * np.ones((nelem, 4)), np.ones((nelem, 4))
    sigy0[mesh.elementgrid[:2, :2].ravel(), :] = 100.0
    if on_gpu:
        vector = gf.VectorPartitionedTyings(conn, dofs, ty.Cdu, ty.Cdp, ty.Cdi)
        elem = gf.Element.Quad4.QuadraturePlanar(vector.AsElement(coor))
        mat = gf.LinearHardening2d(kappa, mu, sigy0, H)
        K = gf.MatrixPartitionedTyings(conn, dofs, ty.Cdu, ty.Cdp)
        solver = gf.MatrixPartitionedTyingsSolver(tol=1e-13)
    else:
        ov = oracle.Vector(conn, dofs)
        elem = oracle.Quadrature(QUAD4_PLANAR, ov.AsElement(coor))
        mat = oracle.LinearHardening(kappa, mu, sigy0, H)
    disp = np.zeros_like(coor)
    fext = np.zeros_like(coor)
    nit = []
    for inc in range(nsteps):
        mat.increment()
        for iiter in range(100):
            if on_gpu:
                elem.symGradN_vector(vector.AsElement(disp), mat.Eps)
                mat.refresh()
                fint = host(vector.AssembleNode(elem.Int_gradN_dot_tensor2_dV(mat.Sig)))
                K.assemble(elem.Int_gradN_dot_tensor4_dot_gradNT_dV(mat.C))
            else:
                mat.Eps = elem.SymGradN_vector(ov.AsElement(disp))
                mat.refresh()
                fint = ov.AsNode(ov.AssembleDofs(elem.Int_gradN_dot_tensor2_dV(mat.Sig)))
                rp, cl, val = oracle.matrix_assemble(conn, dofs, elem.Int_gradN_dot_tensor4_dot_gradNT_dV(mat.C))
                Kd = np.zeros((rp.size - 1,) * 2)
                for r in range(rp.size - 1):
                    Kd[r, cl[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
            fres = fext - fint
            if iiter > 0:
                if on_gpu:
                    Fext, Fint = host(vector.AsDofs_i(fext)), host(vector.AsDofs_i(fint))
                else:
                    fe_d, fi_d = ov.AsDofs(fext), ov.AsDofs(fint)
                    Fext = fe_d[: ty.nni] + ty.Cdi.T @ fe_d[ty.nni:]
                    Fint = fi_d[: ty.nni] + ty.Cdi.T @ fi_d[ty.nni:]
                Fext[ty.nnu:ty.nni] = Fint[ty.nnu:ty.nni]  # copy_p
                nfres, nfext = np.sum(np.abs(Fext - Fint)), np.sum(np.abs(Fext))
                res = nfres / nfext if nfext else nfres
                if res < 1e-5:
                    nit.append(iiter)
                    break
            du = np.zeros_like(coor)
            if iiter == 0:
                du[control_nodes[0], 1] = dgamma
            if on_gpu:
                solver.solve(K, fres, du)
            else:
                du = ov.AsNode(dense_solve(Kd, ty.Cdu, ty.Cdp, ty.nnu, ty.nnp, ov.AsDofs(fres), ov.AsDofs(du)))
            disp += du
        else:
            raise AssertionError("Newton did not converge")
    epsp = host(mat.epsp) if on_gpu else mat.epsp
    return disp, epsp, nit


def test_periodic_elastoplastic_example(gf):
    d_gpu, ep_gpu, it_gpu = run_example(gf, 6, True)
    d_cpu, ep_cpu, it_cpu = run_example(gf, 6, False)
    assert ep_cpu.max() > 0.0  # the soft matrix yields (the hard corner does not)
    assert it_gpu == it_cpu  # the same Newton path
    # the example stops Newton at a 1e-5 relative residual, so the two runs differ by the linear
    # solvers' rounding (CG to 1e-13 against a direct solve) carried through that last iterate
    # (measured 2.5e-8 relative)
    assert np.max(np.abs(d_gpu - d_cpu)) <= 1e-7 * np.max(np.abs(d_cpu))
    assert np.max(np.abs(ep_gpu - ep_cpu)) <= 1e-5 * np.max(ep_cpu)
