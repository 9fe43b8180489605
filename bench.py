"""Benchmark: Quad4 explicit dynamics on Mesh::Quad4::Regular(2048, 2048), plane strain, 2x2 Gauss,
per-element log-normal shear modulus (sigma 0.3), K = 10, lumped mass (density 1), bottom edge
fixed, top edge driven at x-velocity 1e-4, dt = 0.01.  One step = one velocity-Verlet step over the
whole mesh (fused kernel).  Metric: DOF-updates/s = ndof * steps / wall time.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--nx 2048] [--no-cpu] [--no-blocks]

Prints one JSON line.  Inputs are synthetic (SplitMix64/Box-Muller generator, seed 2), resident
in HBM before the timed region.  Beside the headline (FP32 element body) the line carries the FP64
reference-order and FP64-fast legs, their errors over 1000 steps, the roofline against the datasheet
peak and against a measured copy ceiling, the general-mesh path, the one-thread and all-cores CPU
baselines, and the metric's second half: learned vs plain PCG on the periodic Quad4 512^2 cell
(pcg_learned) and the Hex8 128^3 element stiffness on MFMA (hex8_stiffness).  Other workloads:
--workload hex8-stiffness | pcg-periodic | pcg-learned | plastic-refresh | newton-q4 | statics-q4 |
periodic-j2 | quad-api.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E datasheet peak


def canonical_bytes_per_step(nx, ny):
    """Compulsory HBM bytes of one fused step: u, v, a read + written (48 ndof), inverse mass
    (8 ndof), connectivity int32 (16 E), per-element G FP32 (4 E), nodal coordinates (16 N)."""
    nnode = (nx + 1) * (ny + 1)
    ndof = 2 * nnode
    nelem = nx * ny
    return 64 * ndof + 20 * nelem


def setup(nx, ny, seed=2):
    import torch

    import goosefem_amd as gf

    mesh = gf.Mesh.Quad4.Regular(nx, ny)
    coor, conn, dofs = mesh.coor, mesh.conn, mesh.dofs
    vec = gf.Vector(conn, dofs)
    x = vec.AsElement(coor)
    nodal = gf.Element.Quad4.QuadraturePlanar(x, gf.Element.Quad4.Nodal.xi(), gf.Element.Quad4.Nodal.w())
    M = gf.MatrixDiagonal(vec, None)
    M.assemble(nodal.Int_N_scalar_NT_dV(torch.ones((mesh.nelem, 4), dtype=torch.float64, device="cuda")))
    inv = M.Solve(torch.ones(vec.ndof, dtype=torch.float64, device="cuda"))
    del nodal, x
    bot = mesh.nodesBottomEdge
    top = mesh.nodesTopEdge
    fixed = np.concatenate((dofs[bot].ravel(), dofs[top].ravel()))
    inv[torch.from_numpy(fixed).cuda()] = 0.0
    G = np.exp(0.3 * gf.randn(seed, mesh.nelem)).astype(np.float32)
    K = np.full(mesh.nelem, 10.0, dtype=np.float32)
    u = np.zeros(vec.ndof)
    v = np.zeros(vec.ndof)
    v[dofs[top, 0]] = 1e-4
    a = np.zeros(vec.ndof)
    X = gf.ExplicitQuad4(vec, coor, K, G, inv, 0.01)
    X.set_state(u, v, a)
    torch.cuda.synchronize()
    return dict(gf=gf, mesh=mesh, vec=vec, X=X, inv=inv.cpu().numpy(), K=K, G=G, u0=u, v0=v, a0=a,
                conn=conn, dofs=dofs, coor=coor)


def time_steps(X, nsteps, precision, gf):
    """Wall time and average kernel time (HIP events on the launch stream) of nsteps steps."""
    import torch

    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(s)
    X.step(nsteps, precision)
    e1.record(s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return t1 - t0, e0.elapsed_time(e1) * 1e-3


def cpu_baseline(S, nx, ny, nsteps):
    """The oracle's FP64 restatement of the reference loop (single thread), same workload."""
    import oracle

    ov = oracle.Vector(S["conn"], S["dofs"])
    q = oracle.Quadrature(oracle.QUAD4_PLANAR, ov.AsElement(S["coor"]))
    u, v, a = S["u0"].copy(), S["v0"].copy(), S["a0"].copy()
    t0 = time.perf_counter()
    oracle.vv_steps(nsteps, oracle.QUAD4_PLANAR, S["conn"], S["dofs"], q.dNx, q.dV(), S["K"].astype(np.float64),
                    S["G"].astype(np.float64), S["inv"], 0.01, u, v, a)
    dt = time.perf_counter() - t0
    return ov.ndof * nsteps / dt, dt


def load_traffic(nx, name="explicit_vv_quad4_2048_traffic.json", size_key="nx", key="traffic_bytes_per_launch"):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if it matches this size."""
    p = os.path.join(ROOT, "profiles", name)
    try:
        d = json.load(open(p))
        if d.get(size_key) == nx:
            return d.get(key)
    except (OSError, ValueError):
        pass
    return None


def _timed(fn, reps):
    """Average seconds per call of fn() between HIP events on the launch stream (after 1 warm-up)."""
    import torch

    fn()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / reps


def bench_hex8(args):
    """Hex8 Regular n^3, per-point isotropic tangent (K = 10, G ~ U[0.5, 1.5]): element stiffness in
    FP64 vs BF16x3 MFMA, CSR assembly, assembled Dot vs matrix-free sum_e K_e u_e."""
    print(json.dumps(hex8_stiffness(args.hex_n, cpu=not args.no_cpu, csr=True)))


# one v_mfma_f32_32x32x16_bf16 keeps its SIMD's matrix core busy 32 cycles; 15 per element
MFMA_PER_ELEMENT, MFMA_CYCLES, SIMDS, CLOCK_HZ = 15, 32, 1024, 2.4e9


def hex8_stiffness(n, cpu=True, csr=True, cpu_seconds=10.0):
    """The Hex8 element-stiffness workload as a dict (see bench_hex8)."""
    import torch

    import goosefem_amd as gf

    t0 = time.perf_counter()
    mesh = gf.Mesh.Hex8.Regular(n, n, n)
    coor, conn, dofs = mesh.coor, mesh.conn, mesh.dofs
    vec = gf.Vector(conn, dofs)
    quad = gf.Element.Hex8.Quadrature(vec.AsElement(coor))
    E = mesh.nelem
    z = torch.from_numpy(gf.randn(3, E * 8)).cuda()
    Gq = (0.5 + 0.5 * (1.0 + torch.erf(z / np.sqrt(2.0)))).reshape(E, 8)  # U[0.5, 1.5]
    Kq = torch.full((E, 8), 10.0, dtype=torch.float64, device="cuda")
    C = gf.elastic_tangent(Kq, Gq)
    C32 = gf.elastic_tangent(Kq, Gq, dtype=torch.float32)
    k64 = torch.empty((E, 24, 24), dtype=torch.float64, device="cuda")
    kbf = torch.empty_like(k64)
    k32 = torch.empty((E, 24, 24), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    t_tan = _timed(lambda: gf.elastic_tangent(Kq, Gq), 3)
    t_tan32 = _timed(lambda: gf.elastic_tangent(Kq, Gq, dtype=torch.float32), 3)
    del z, Gq, Kq
    t64 = _timed(lambda: quad.Int_gradN_dot_tensor4_dot_gradNT_dV(C, gf.F64, out=k64), 3)
    tbf = _timed(lambda: quad.Int_gradN_dot_tensor4_dot_gradNT_dV(C, gf.BF16X3, out=kbf), 5)
    t32 = _timed(lambda: quad.Int_gradN_dot_tensor4_dot_gradNT_dV(C32, gf.BF16X3, out=k32), 5)
    err_ke = float(torch.linalg.vector_norm(kbf - k64) / torch.linalg.vector_norm(k64))
    err_ke32 = float(torch.linalg.vector_norm(k32.double() - k64) / torch.linalg.vector_norm(k64))
    del C, C32
    bytes_el = 11392  # FP64 API bytes per element: C 5184 + GradN 1536 + dV 64 + K_e 4608
    flops_el = 69120  # Voigt B^T (D dV) B, 8 points
    ach = bytes_el * E / tbf / 1e9
    res = {
        "metric": "Hex8 element stiffness (BF16x3 MFMA) throughput",
        "value": E / tbf,
        "unit": "elements/s",
        "config": {"workload": f"Hex8 Regular {n}^3 element stiffness" + (" + CSR assembly + K.u" if csr else ""),
                   "nelem": E, "ndof": vec.ndof},
        "dtype": "bf16x3 (hi/lo split) mfma, fp32 accumulate; fp64 inputs/outputs",
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": load_traffic(n, "hex8_stiffness_128_traffic.json", "n"), "kernel": "k_stiffness_bf16x3",
                     "algorithmic_bytes_per_launch": bytes_el * E, "avg_launch_us": tbf * 1e6,
                     "tflops_voigt": flops_el * E / tbf / 1e12},
        # matrix cores busy: MFMA issue cycles of the launch over the chip's SIMD cycles in its duration
        "mfma_busy_est": MFMA_PER_ELEMENT * MFMA_CYCLES * E / (SIMDS * CLOCK_HZ * tbf),
        "fp64_stiffness": {"avg_launch_us": t64 * 1e6, "achieved_GBs": bytes_el * E / t64 / 1e9},
        # FP32 tangent in, FP32 K_e out (the kernel's own operand and accumulator precision):
        # C 2592 + GradN 1536 + dV 64 + K_e 2304 = 6496 B per element
        "fp32_io": {"avg_launch_us": t32 * 1e6, "elements_per_s": E / t32,
                    "traffic": load_traffic(n, "hex8_stiffness_128_traffic.json", "n", "traffic_bytes_per_launch_f32io"),
                    "achieved_GBs": 6496 * E / t32 / 1e9, "frac": 6496 * E / t32 / 1e9 / HBM_PEAK_GBS,
                    "mfma_busy_est": MFMA_PER_ELEMENT * MFMA_CYCLES * E / (SIMDS * CLOCK_HZ * t32),
                    "Ke_rel_vs_fp64": err_ke32},
        "tangent_ms": {"f64": t_tan * 1e3, "f32": t_tan32 * 1e3,
                       "f64_GBs": E * 8 * 81 * 8 / t_tan / 1e9, "f32_GBs": E * 8 * 81 * 4 / t_tan32 / 1e9},
        "error_vs_fp64": {"Ke_rel": err_ke},
        "setup_s": t_setup,
    }
    if csr:
        t0 = time.perf_counter()
        A = gf.Matrix(vec, None)
        t_pattern = time.perf_counter() - t0
        t_asm = _timed(lambda: A.assemble(k64), 3)
        u = torch.from_numpy(gf.randn(4, vec.ndof)).cuda()
        Ku = torch.empty_like(u)
        t_dot = _timed(lambda: A.Dot(u, out=Ku), 10)

        def matrix_free(K):
            return vec.AssembleDofs(gf.elemmat_dot(K, vec.AsElement(u)))

        t_mf = _timed(lambda: matrix_free(k64), 3)
        err_mf = float(torch.linalg.vector_norm(matrix_free(k64) - Ku) / torch.linalg.vector_norm(Ku))
        err_ku_bf = float(torch.linalg.vector_norm(matrix_free(kbf) - Ku) / torch.linalg.vector_norm(Ku))
        # the same products from the FP32 element matrices (FP64 sums)
        t_mf32 = _timed(lambda: matrix_free(k32), 3)
        err_ku32 = float(torch.linalg.vector_norm(matrix_free(k32) - Ku) / torch.linalg.vector_norm(Ku))
        t_asm32 = _timed(lambda: A.assemble(k32), 3)
        A.assemble(k64)
        res["config"]["nnz"] = A.nnz
        res["error_vs_fp64"]["Ku_rel"] = err_ku_bf
        res["fp32_io"].update({"assemble_ms": t_asm32 * 1e3, "assemble_GBs": (E * 2304 + A.nnz * 8) / t_asm32 / 1e9,
                               "matrix_free_ms": t_mf32 * 1e3, "Ku_rel_vs_fp64": err_ku32})
        res["csr"] = {"pattern_build_s": t_pattern, "assemble_ms": t_asm * 1e3,
                      "assemble_GBs": (E * 4608 + A.nnz * 8) / t_asm / 1e9,
                      "dot_ms": t_dot * 1e3, "dot_GBs": (A.nnz * 12 + vec.ndof * 16) / t_dot / 1e9,
                      "matrix_free_ms": t_mf * 1e3, "matrix_free_vs_dot_rel": err_mf}
    if cpu:
        import oracle

        # bounded CPU sample: chunks of 50,000 elements of the same mesh until about cpu_seconds of work
        m, dt, e0 = 50000, 0.0, 0
        Gs = 0.5 + np.random.default_rng(0).random((m, 8))
        Cs = oracle.elastic_tangent(np.full((m, 8), 10.0), Gs)
        while dt < cpu_seconds and e0 + m <= conn.shape[0]:
            ov = oracle.Vector(conn[e0:e0 + m], oracle.dofs(coor.shape[0], 3))
            oq = oracle.Quadrature(oracle.HEX8, ov.AsElement(coor))
            t0 = time.perf_counter()
            oq.Int_gradN_dot_tensor4_dot_gradNT_dV(Cs)
            dt += time.perf_counter() - t0
            e0 += m
        res["cpu_baseline"] = {"value": e0 / dt, "unit": "elements/s", "cores": 1, "kind": "port",
                               "sample": f"{e0} elements of the same mesh, FP64 oracle reference loop, one thread, "
                                         f"{dt:.1f} s"}
    return res


def periodic_setup(n, hex8=False, seed=4, sigma=0.4, spheres=False):
    """Periodic Regular mesh with Tyings::Control + Tyings::Periodic as in
    docs/examples/statics/periodic_elastoplastic.py:28-47; per-element G log-normal, K = 10;
    with spheres (Hex8) G is multiplied by 10 inside periodic spheres of radius 4 elements filling
    20% of the element centres."""
    import goosefem_amd as gf

    mesh = gf.Mesh.Hex8.Regular(n, n, n) if hex8 else gf.Mesh.Quad4.Regular(n, n)
    coor, conn = mesh.coor, mesh.conn
    ctrl = gf.Tyings.Control(coor, mesh.dofs)
    iip = np.concatenate((ctrl.controlDofs.ravel(), ctrl.dofs[mesh.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, mesh.nodesPeriodic, iip)
    vec = gf.Vector(conn, ty.dofs)
    x = gf.Vector(conn, mesh.dofs).AsElement(coor)
    quad = gf.Element.Hex8.Quadrature(x) if hex8 else gf.Element.Quad4.QuadraturePlanar(x)
    G = np.exp(sigma * gf.randn(seed, mesh.nelem))
    nsph = 0
    if spheres:
        from goosefem_amd.warmstart import stiff_spheres

        mask, nsph = stiff_spheres(n)
        G[mask] *= 10.0
    K = np.full(mesh.nelem, 10.0)
    P = gf.TyingsPCG(vec, quad, K, G, ty)
    return dict(gf=gf, mesh=mesh, ty=ty, vec=vec, quad=quad, G=G, K=K, P=P, conn=conn, coor=coor, nspheres=nsph)


def shear_rhs(S, dgamma, pure=False):
    """b_u = -K'_up x_p for a shear increment: simple shear ctrl(0, 1) = dgamma, or pure shear
    ctrl(0, 1) = ctrl(1, 0) = dgamma."""
    ty, P = S["ty"], S["P"]
    xp = np.zeros(ty.nnp)
    xp[ty.control[0, 1] - ty.nnu] = dgamma
    if pure:
        xp[ty.control[1, 0] - ty.nnu] = dgamma
    return -P.apply(None, xp), xp


def bench_pcg(args):
    """Plain FP64 Jacobi-PCG on the periodic tied operator, zero initial guess, to 1e-10."""
    import torch

    n = args.pcg_n
    S = periodic_setup(n, args.pcg_hex, spheres=args.spheres)
    P, ty = S["P"], S["ty"]
    gamma = 5e-4 if args.spheres else 1e-3
    b, xp = shear_rhs(S, gamma, pure=args.spheres)
    torch.cuda.synchronize()
    P.solve(b, tol=1e-10)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, it, res = P.solve(b, tol=1e-10)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    res_d = {"metric": "PCG time to relative residual 1e-10 (plain, zero guess)", "value": t, "unit": "s",
             "higher_is_better": False,
             "config": {"workload": f"periodic {'Hex8' if args.pcg_hex else 'Quad4'} Regular {n}"
                                    f"{'^3' if args.pcg_hex else '^2'}, log-normal G (sigma 0.4)"
                                    f"{', G x10 in stiff spheres (r 4, 20%)' if args.spheres else ''}, K 10, "
                                    f"{'pure' if args.spheres else 'simple'} shear increment {gamma}",
                        "nnu": ty.nnu, "spheres": S["nspheres"], "structured_operator": P.structured},
             "iterations": it, "true_relres": res, "ms_per_iteration": t / max(it, 1) * 1e3}
    if not args.no_cpu:
        res_d["cpu_baseline"] = pcg_cpu_baseline(S, b, target_s=10.0)
    print(json.dumps(res_d))


def pcg_cpu_baseline(S, b, target_s=10.0):
    """The oracle's FP64 CG on the assembled K' (scipy sparse products) on one thread, a bounded
    number of iterations sized to about target_s seconds; reported per iteration."""
    import scipy.sparse as sp

    import oracle

    ty = S["ty"]
    conn = S["conn"]
    q = oracle.Quadrature(oracle.HEX8 if S["mesh"].ndim == 3 else oracle.QUAD4_PLANAR,
                          oracle.Vector(conn, oracle.dofs(S["coor"].shape[0], S["coor"].shape[1])).AsElement(S["coor"]))
    nip = q.nip
    C = oracle.elastic_tangent(np.repeat(S["K"][:, None], nip, 1), np.repeat(S["G"][:, None], nip, 1))
    Ke = q.Int_gradN_dot_tensor4_dot_gradNT_dV(C)
    rp, cl, vl = oracle.matrix_assemble(conn, ty.dofs[: S["coor"].shape[0]], Ke)
    ndof = ty.nni + ty.nnd
    A = sp.csr_matrix((vl, cl, rp), shape=(ndof, ndof))
    u, d = slice(0, ty.nnu), slice(ty.nni, ndof)
    crp, ccl, cvl, ncu = ty.Cdu_csr  # sparse: the dense form is nnd x nnu
    Cdu = sp.csr_matrix((cvl, ccl, crp), shape=(crp.size - 1, ncu))
    Kp = (A[u, u] + A[u, d] @ Cdu + Cdu.T @ A[d, u] + Cdu.T @ A[d, d] @ Cdu).tocsr()
    Kp.sort_indices()
    bh = b.cpu().numpy()
    args = (Kp.indptr.astype(np.int64), Kp.indices.astype(np.int64), Kp.data, Kp.diagonal().copy(), bh)
    t0 = time.perf_counter()
    oracle.pcg_csr(*args, tol=1e-30, maxiter=20)
    per = (time.perf_counter() - t0) / 20
    iters = int(min(20000, max(50, target_s / per)))
    t0 = time.perf_counter()
    _, it, _ = oracle.pcg_csr(*args, tol=1e-30, maxiter=iters)
    dt = time.perf_counter() - t0
    return {"value": dt / it * 1e3, "unit": "ms/iteration", "cores": 1, "kind": "port",
            "sample": f"{it} CG iterations on the assembled K' ({Kp.nnz} nnz), FP64 oracle, one thread, {dt:.1f} s"}


def bench_pcg_learned(args):
    """Wall time of the FP64 Jacobi-PCG to a true relative residual of 1e-10 on the periodic cell
    (Quad4 512^2 simple shear, or Hex8 96^3 with stiff spheres under pure shear) from the learned
    warm start (network inference included) and from a zero guess, on modulus fields the network
    was not trained on; with --increments N also the quasi-static loop of N load increments."""
    if not os.path.exists(args.ws_weights):
        print(json.dumps({"error": f"no trained weights at {args.ws_weights}; run "
                                   "python -m goosefem_amd.warmstart --out <file> first"}))
        sys.exit(2)
    print(json.dumps(pcg_learned(args.ws_weights, args.ws_tests, cpu=not args.no_cpu, increments=args.increments)))


def pcg_learned(weights, ntests, cpu=False, increments=0, cpu_seconds=10.0):
    """Learned warm start vs plain PCG (see bench_pcg_learned) as a dict."""
    import torch

    import goosefem_amd as gf
    from goosefem_amd import warmstart as ws

    t_setup = time.perf_counter()
    net = ws.WarmStartNet.load(weights)
    n = int(net.meta["n"])
    dim = int(net.meta.get("dim", 2))
    shear = str(net.meta.get("shear", "simple"))
    pb = ws.periodic_problem(n, sigma=float(net.meta["sigma"]), dim=dim, spheres=dim == 3)
    W = ws.PeriodicWarmStart(pb.P, pb.ty, n, net, sigma=pb.sigma, dim=dim, shear=shear)
    P = pb.P
    gamma = 5e-4 if dim == 3 else 1e-3
    Kd = gf.to_device(pb.K)
    zero_grid = torch.zeros(W.nv * dim, dtype=torch.float32, device="cuda")
    t_setup = time.perf_counter() - t_setup

    def affine_guess():
        x = torch.zeros(pb.ty.nnu, dtype=torch.float64, device="cuda")
        gf._call("gf_grid_to_dofs", gf._i64(W.nv * dim), gf._i32(dim), gf._dp(zero_grid), gf._dp(W.idx),
                 gf._dp(W.affine), gf._f64(1.0), gf._f64(gamma), gf._dp(x), gf._stream())
        return x

    def clock(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    rows = []
    for k in range(ntests + 1):
        seed = 900000 + k  # training fields use seeds 1000 + i
        g = ws.sphere_field(n, seed, pb.sigma) if dim == 3 else np.exp(pb.sigma * gf.randn(seed, n * n))
        G = gf.to_device(g)
        P.set_moduli(Kd, G)
        b = W.rhs(gamma)
        t_plain, (_, it0, res0) = clock(lambda: P.solve(b, tol=1e-10))
        t_inf, _ = clock(lambda: W.predict(G, gamma))
        t_learned, (_, it1, res1) = clock(lambda: P.solve(b, x0=W.predict(G, gamma), tol=1e-10))
        start = P.start_used
        _, it2, _ = P.solve(b, x0=affine_guess(), tol=1e-10)
        if k == 0:
            continue  # first field: warm-up of kernels and buffers
        rows.append(dict(seed=seed, plain_s=t_plain, learned_s=t_learned, inference_s=t_inf, plain_iterations=it0,
                         learned_iterations=it1, affine_only_iterations=it2, plain_relres=res0, learned_relres=res1,
                         learned_start_kept=start == gf.GF_START_GIVEN))
    mean = lambda key: float(np.mean([r[key] for r in rows]))  # noqa: E731
    m = net.meta
    if dim == 3:
        workload = (f"periodic Hex8 Regular {n}^3, G = exp({pb.sigma} z) x10 in stiff spheres (r 4, 20%), K 10, "
                    f"pure shear increment {gamma}")
    else:
        workload = f"periodic Quad4 Regular {n}^2 plane strain, G = exp({pb.sigma} z), K 10, simple shear increment {gamma}"
    res = {
        "metric": "PCG wall time to true relative residual 1e-10, learned warm start (inference included)",
        "value": mean("learned_s"), "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "plain_s": mean("plain_s"), "learned_s": mean("learned_s"),
        "speedup_vs_plain": mean("plain_s") / mean("learned_s"),
        "iterations": {"plain": mean("plain_iterations"), "learned": mean("learned_iterations"),
                       "affine_only": mean("affine_only_iterations")},
        "inference_ms": mean("inference_s") * 1e3,
        "max_true_relres": max(max(r["plain_relres"], r["learned_relres"]) for r in rows),
        "dtype": "bf16 mfma network (fp32 accumulate), fp64 pcg",
        "data": "synthetic modulus fields (held-out seeds)",
        "config": {"workload": workload, "nnu": pb.ty.nnu, "test_fields": len(rows), "channels": net.ch,
                   "parameters": net.nparams(), "structured_operator": P.structured},
        "training": {"samples": int(m["samples"]), "epochs": int(m["epochs"]), "batch": int(m["batch"]),
                     "data_generation_s": float(m["data_seconds"]), "training_s": float(m["train_seconds"]),
                     "final_mse": float(m["final_mse"]),
                     "holdout_rel_error": float(m.get("holdout_rel_error", float("nan")))},
        "setup_s": t_setup,
        "per_field": rows,
    }
    if increments > 0:
        # the quasi-static loop of the reference's periodic example (periodic_elastoplastic.py:89-146)
        # for a linear material: every increment solves K' du_u = -K'_up du_p for the same du_p, so
        # the previous increment's du_u is an exact start; four start strategies, each to 1e-10
        g = np.exp(pb.sigma * gf.randn(900001, n * n)) if dim == 2 else ws.sphere_field(n, 900001, pb.sigma)
        G = gf.to_device(g)
        P.set_moduli(Kd, G)
        b = W.rhs(gamma)
        loops = {}
        for name in ("plain_zero", "plain_previous", "learned", "learned_then_previous"):
            du = None
            its = []
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for inc in range(increments):
                if name == "plain_zero" or (name == "plain_previous" and inc == 0):
                    x0 = None
                elif name == "learned" or inc == 0:
                    x0 = W.predict(G, gamma)
                else:
                    x0 = du
                du, it, r = P.solve(b, x0=x0, tol=1e-10)
                its.append(it)
            torch.cuda.synchronize()
            loops[name] = {"s": time.perf_counter() - t0, "iterations_total": int(np.sum(its)),
                           "iterations_first": int(its[0]), "iterations_rest_mean": float(np.mean(its[1:]))
                           if len(its) > 1 else 0.0}
        res["quasi_static"] = {"increments": increments, "increment": gamma, "field_seed": 900001, **loops,
                               "speedup_learned_vs_plain_zero": loops["plain_zero"]["s"] / loops["learned"]["s"],
                               "speedup_learned_then_previous_vs_plain_previous":
                                   loops["plain_previous"]["s"] / loops["learned_then_previous"]["s"]}
    if cpu:
        S = dict(ty=pb.ty, conn=pb.mesh.conn, coor=pb.mesh.coor, K=pb.K, G=np.asarray(g, dtype=np.float64),
                 mesh=pb.mesh)
        res["cpu_baseline"] = pcg_cpu_baseline(S, W.rhs(gamma), target_s=cpu_seconds)
    return res


def bench_plastic(args):
    """J2 linear-hardening refresh (stress, consistent tangent, plastic state) at every quadrature
    point of a Hex8 Regular n^3 mesh (8 points per element): the per-point step of the statics
    examples' Newton loop.  Strains are random symmetric tensors, a third of them scaled to stay
    elastic; moduli K = 10, G = 1, sigy0 = 0.05, H = 0.4."""
    import torch

    import goosefem_amd as gf

    n = args.hex_n
    nq = n * n * n * 8
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    E = torch.randn((nq, 3, 3), dtype=torch.float64, device="cuda", generator=g) * 0.05
    E = 0.5 * (E + E.transpose(1, 2))
    E[::3] *= 0.01
    mat = gf.LinearHardening(torch.full((nq,), 10.0, dtype=torch.float64, device="cuda"), 1.0, 0.05, 0.4)
    mat.Eps = E
    del E
    torch.cuda.synchronize()
    t_full = _timed(lambda: mat.refresh(True), 10)
    t_sig = _timed(lambda: mat.refresh(False), 10)
    frac_plastic = float((mat.epsp > 0).double().mean())
    # algorithmic bytes per point: read Eps 72 + Epsp_t 72 + (K, G, sigy0, H) 32 + epsp_t 8; write
    # Sig 72 + Epsp 72 + epsp 8, and C 648 with the tangent
    b_full, b_sig = 984, 336
    ach = b_full * nq / t_full / 1e9
    res = {
        "metric": "J2 linear-hardening refresh throughput (stress + consistent tangent)",
        "value": nq / t_full,
        "unit": "points/s",
        "config": {"workload": f"Hex8 Regular {n}^3 quadrature points (x8), J2 linear hardening", "points": nq,
                   "plastic_fraction": frac_plastic},
        "dtype": "f64",
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "kernel": "k_plastic_refresh",
                     "traffic": load_traffic(n, "plastic_refresh_hex8_128_traffic.json", "n"),
                     "algorithmic_bytes_per_launch": b_full * nq, "avg_launch_us": t_full * 1e6},
        "stress_only": {"avg_launch_us": t_sig * 1e6, "points_per_s": nq / t_sig,
                        "achieved_GBs": b_sig * nq / t_sig / 1e9, "frac": b_sig * nq / t_sig / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_cpu:
        import oracle

        rng = np.random.default_rng(0)
        m = 1 << 20
        Es = rng.standard_normal((m, 3, 3)) * 0.05
        Es = 0.5 * (Es + Es.transpose(0, 2, 1))
        Es[::3] *= 0.01
        z9, z1 = np.zeros((m, 3, 3)), np.zeros(m)
        done, dt = 0, 0.0
        while dt < 10.0:
            t0 = time.perf_counter()
            oracle.plastic_refresh(Es, 10.0, 1.0, 0.05, 0.4, z9, z1)
            dt += time.perf_counter() - t0
            done += m
        res["cpu_baseline"] = {"value": done / dt, "unit": "points/s", "cores": 1, "kind": "port",
                               "sample": f"{done} points (batches of {m}), FP64 oracle restatement, one thread, "
                                         f"{dt:.1f} s"}
    print(json.dumps(res))


def bench_newton(args):
    """One pass of the statics Newton iteration with J2 linear hardening on a Quad4 plane-strain
    Regular n^2 mesh (4 points per element): strain, return map, element force and element
    stiffness.  Fused (gf_plastic_newton_q4: Eps and C stay on chip) against the unfused calls
    (symGradN_vector, refresh, Int_gradN_dot_tensor2_dV, Int_gradN_dot_tensor4_dot_gradNT_dV)."""
    import torch

    import goosefem_amd as gf

    n = args.newton_n
    mesh = gf.Mesh.Quad4.Regular(n, n)
    vec = gf.Vector(mesh.conn, mesh.dofs)
    quad = gf.Element.Quad4.QuadraturePlanar(vec.AsElement(mesh.coor))
    ne = mesh.nelem
    u = torch.from_numpy(gf.randn(11, 2 * mesh.nnode, 0.0, 2e-3)).cuda().reshape(mesh.nnode, 2)
    ue = vec.AsElement(u)
    Kq = torch.full((ne, 4), 10.0, dtype=torch.float64, device="cuda")
    fused = gf.LinearHardening(Kq, 1.0, 0.01, 0.4)
    plain = gf.LinearHardening(Kq, 1.0, 0.01, 0.4)
    fe = torch.empty((ne, 4, 2), dtype=torch.float64, device="cuda")
    Ke = torch.empty((ne, 8, 8), dtype=torch.float64, device="cuda")
    fe2, Ke2 = torch.empty_like(fe), torch.empty_like(Ke)
    torch.cuda.synchronize()
    t_fused = _timed(lambda: fused.newton_q4(quad, ue, fe, Ke), 10)

    def unfused():
        quad.symGradN_vector(ue, plain.Eps)
        plain.refresh()
        quad.Int_gradN_dot_tensor2_dV(plain.Sig, out=fe2)
        quad.Int_gradN_dot_tensor4_dot_gradNT_dV(plain.C, out=Ke2)

    t_plain = _timed(unfused, 5)
    same = bool(torch.equal(fe, fe2) and torch.equal(Ke, Ke2))
    frac_plastic = float((fused.epsp > 0).double().mean())
    # algorithmic bytes per element (FP64): reads dNx 256, dV 32, u_e 64, (K, G, sigy0, H) 128,
    # Epsp_t 288, epsp_t 32; writes Epsp 288, epsp 32, f_e 64, K_e 512
    b_el = 1696
    b_plain = 8576  # the unfused chain: Eps, Sig and C written and read back
    ach = b_el * ne / t_fused / 1e9
    res = {
        "metric": "J2 Newton pass throughput (fused strain + return map + f_e + K_e)",
        "value": ne / t_fused,
        "unit": "elements/s",
        "config": {"workload": f"Quad4 plane strain Regular {n}^2, J2 linear hardening, 4 points per element",
                   "nelem": ne, "plastic_fraction": frac_plastic},
        "dtype": "f64",
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "kernel": "k_plastic_newton_q4", "traffic": load_traffic(n, "newton_q4_2048_traffic.json", "n"),
                     "algorithmic_bytes_per_launch": b_el * ne, "avg_launch_us": t_fused * 1e6},
        "unfused": {"ms": t_plain * 1e3, "algorithmic_GB": b_plain * ne / 1e9, "speedup_fused": t_plain / t_fused,
                    "bit_identical": same},
    }
    print(json.dumps(res))


def bench_statics(args):
    """One load increment of the statics example on the device: a Quad4 plane-strain n^2 plate, J2
    linear hardening, bottom edge fixed, top edge sheared; Newton iterations of the fused element
    pass, MatrixPartitioned assembly (CSR + the four blocks) and the partitioned Jacobi-PCG solve."""
    import torch

    import goosefem_amd as gf

    n = args.statics_n
    mesh = gf.Mesh.Quad4.Regular(n, n)
    dofs = mesh.dofs
    iip = np.concatenate((dofs[mesh.nodesBottomEdge].ravel(), dofs[mesh.nodesTopEdge].ravel()))
    vector = gf.VectorPartitioned(mesh.conn, dofs, iip)
    quad = gf.Element.Quad4.QuadraturePlanar(vector.AsElement(mesh.coor))
    ne = mesh.nelem
    mat = gf.LinearHardening(torch.full((ne, 4), 10.0, dtype=torch.float64, device="cuda"), 1.0, 0.005, 0.4)
    t0 = time.perf_counter()
    A = gf.MatrixPartitioned(vector)
    t_pattern = time.perf_counter() - t0
    solver = gf.MatrixPartitionedSolver(tol=1e-10)
    disp = torch.zeros((mesh.nnode, 2), dtype=torch.float64, device="cuda")
    fe = torch.empty((ne, 4, 2), dtype=torch.float64, device="cuda")
    Ke = torch.empty((ne, 8, 8), dtype=torch.float64, device="cuda")
    t_el = t_asm = t_solve = 0.0
    its, newton, resid = [], 0, []
    mat.increment()
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for it in range(20):
        ta = time.perf_counter()
        mat.newton_q4(quad, vector.AsElement(disp), fe, Ke)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        A.assemble(Ke)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        fint = vector.AssembleNode(fe)
        fres = -fint
        if it > 0:
            fres[torch.from_numpy(vector.dofs_is_p()).cuda()] = 0.0
            rel = float(fres.abs().sum()) / float(fint.abs().sum())
            resid.append(rel)
            if rel < 1e-8:
                break
        du = torch.zeros_like(disp)
        if it == 0:
            du[torch.from_numpy(mesh.nodesTopEdge).cuda(), 0] = 0.005 * n
        solver.solve(A, fres, du)
        torch.cuda.synchronize()
        td = time.perf_counter()
        disp += du
        t_el += tb - ta
        t_asm += tc - tb
        t_solve += td - tc
        its.append(solver.iterations)
        newton += 1
    t_all = time.perf_counter() - t_all
    nnz_uu = A.nnz_blocks[0]
    nnu = vector.nnu
    cg_its = sum(its)
    t_it = t_solve / max(cg_its, 1)
    # per CG iteration: A_uu p (12 B/nnz + p, q), the update (p, q, diag, x, r read; x, r, z
    # written) and the direction (z, p read; p written): 12 nnz + 15 * 8 nnu bytes
    b_it = 12 * nnz_uu + 120 * nnu
    ach = b_it / t_it / 1e9
    res = {
        "metric": "Statics load increment (J2 plate, Newton + partitioned PCG) wall time",
        "value": t_all,
        "unit": "s",
        "higher_is_better": False,
        "config": {"workload": f"Quad4 plane strain Regular {n}^2, J2 linear hardening, bottom fixed, top sheared",
                   "nelem": ne, "nnu": nnu, "nnz_uu": nnz_uu, "newton_iterations": newton, "cg_iterations": its,
                   "newton_residuals": resid,
                   "plastic_fraction": float((mat.epsp > 0).double().mean())},
        "dtype": "f64",
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "kernel": "partitioned Jacobi-PCG iteration (k_cg_apply_dir + k_cg_update)",
                     "traffic": None, "algorithmic_bytes_per_launch": b_it, "avg_launch_us": t_it * 1e6},
        "breakdown_s": {"fused_element_pass": t_el, "matrix_assemble": t_asm, "solve": t_solve,
                        "pattern_build": t_pattern},
    }
    print(json.dumps(res))


def bench_periodic_j2(args):
    """The reference's periodic elasto-plastic example (periodic_elastoplastic.py) at n^2 elements:
    J2 linear hardening (hard 2x2-element corner), shear increments of 0.01 through the control
    node, Newton with the assembled tangent and the tied solve (MatrixPartitionedTyings +
    MatrixPartitionedTyingsSolver: matrix-free condensed CG), all on the device."""
    import torch

    import goosefem_amd as gf

    n = args.periodic_n
    mesh = gf.Mesh.Quad4.Regular(n, n)
    control = gf.Tyings.Control(mesh.coor, mesh.dofs)
    coor, dofs = control.coor, control.dofs
    cdofs, cnodes = control.controlDofs, control.controlNodes
    iip = np.concatenate((cdofs.ravel(), dofs[mesh.nodesOrigin, :].ravel()))
    ty = gf.Tyings.Periodic(coor, dofs, cdofs, mesh.nodesPeriodic, iip)
    dofs = ty.dofs
    Cdu, Cdp, Cdi = ty.Cdu_csr, ty.Cdp_csr, ty.Cdi_csr
    vector = gf.VectorPartitionedTyings(mesh.conn, dofs, Cdu, Cdp, Cdi)
    elem = gf.Element.Quad4.QuadraturePlanar(vector.AsElement(coor))
    ne = mesh.nelem
    sy = np.full((ne, 4), 0.05)
    sy[mesh.elementgrid[:2, :2].ravel(), :] = 100.0
    mat = gf.LinearHardening(np.ones((ne, 4)), 1.0, sy, 1.0)
    K = gf.MatrixPartitionedTyings(mesh.conn, dofs, Cdu, Cdp)
    solver = gf.MatrixPartitionedTyingsSolver(tol=1e-10)
    disp = torch.zeros((coor.shape[0], 2), dtype=torch.float64, device="cuda")
    fext = torch.zeros_like(disp)
    fe = torch.empty((ne, 4, 2), dtype=torch.float64, device="cuda")
    Ke = torch.empty((ne, 8, 8), dtype=torch.float64, device="cuda")
    newton, cg, t_solve = [], [], 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for inc in range(args.periodic_steps):
        mat.increment()
        for it in range(50):
            mat.newton_q4(elem, vector.AsElement(disp), fe, Ke)
            fint = vector.AssembleNode(fe)
            K.assemble(Ke)
            fres = fext - fint
            if it > 0:
                Fext, Fint = vector.AsDofs_i(fext), vector.AsDofs_i(fint)
                Fext[ty.nnu:ty.nni] = Fint[ty.nnu:ty.nni]
                nfext = float(Fext.abs().sum())
                res = float((Fext - Fint).abs().sum()) / (nfext if nfext else 1.0)
                if res < 1e-5:
                    newton.append(it)
                    break
            du = torch.zeros_like(disp)
            if it == 0:
                du[int(cnodes[0]), 1] = 0.01
            torch.cuda.synchronize()
            ts = time.perf_counter()
            solver.solve(K, fres, du)
            torch.cuda.synchronize()
            t_solve += time.perf_counter() - ts
            cg.append(solver.iterations)
            disp += du
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res = {
        "metric": "Periodic elasto-plastic example (Newton + tied solve) wall time",
        "value": wall,
        "unit": "s",
        "higher_is_better": False,
        "config": {"workload": f"periodic_elastoplastic.py on Quad4 Regular {n}^2, {args.periodic_steps} shear increments of 0.01",
                   "nelem": ne, "nnu": ty.nnu, "newton_iterations": newton, "cg_iterations": cg,
                   "plastic_fraction": float((mat.epsp > 0).double().mean())},
        "dtype": "f64",
        "us_per_cg_iteration": t_solve / max(sum(cg), 1) * 1e6,
        "solve_s": t_solve,
    }
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-blocks", action="store_true",
                    help="skip the learned-PCG (512^2 periodic) and Hex8 stiffness (128^3) blocks")
    ap.add_argument("--no-general", action="store_true", help="skip the general-mesh explicit row")
    ap.add_argument("--no-errors", action="store_true", help="skip the FP32 / FP64-fast error trajectories (profiling)")
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--workload", choices=["quad4-explicit", "hex8-stiffness", "pcg-periodic", "pcg-learned",
                                           "plastic-refresh", "newton-q4", "statics-q4", "periodic-j2", "quad-api"],
                    default="quad4-explicit")
    ap.add_argument("--pcg-n", type=int, default=512)
    ap.add_argument("--pcg-hex", action="store_true", help="periodic Hex8 n^3 instead of Quad4 n^2")
    ap.add_argument("--spheres", action="store_true",
                    help="stiff periodic spheres (G x10, radius 4, 20%% of elements) and pure shear 5e-4")
    ap.add_argument("--hex-n", type=int, default=128)
    ap.add_argument("--newton-n", type=int, default=2048)
    ap.add_argument("--statics-n", type=int, default=256)
    ap.add_argument("--periodic-n", type=int, default=128)
    ap.add_argument("--periodic-steps", type=int, default=4)
    ap.add_argument("--ws-weights", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "goosefem_amd",
                                                         "data", "warmstart_quad4_512.npz"))
    ap.add_argument("--ws-tests", type=int, default=4, help="held-out modulus fields timed")
    ap.add_argument("--increments", type=int, default=0,
                    help="pcg-learned: also run the quasi-static loop of this many load increments")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (PMC counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; the path shards by mesh slabs "
                                   "(halo exchange of one node row) but no multi-GPU run is provided"}))
        sys.exit(2)
    import torch

    if args.copy_calibration:
        # plain device-to-device copy of a known size: its FETCH_SIZE / WRITE_SIZE counters give the
        # scale between counter units and bytes
        n = (1 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))
        return

    if args.workload == "hex8-stiffness":
        bench_hex8(args)
        return
    if args.workload == "pcg-periodic":
        bench_pcg(args)
        return
    if args.workload == "pcg-learned":
        bench_pcg_learned(args)
        return
    if args.workload == "plastic-refresh":
        bench_plastic(args)
        return
    if args.workload == "newton-q4":
        bench_newton(args)
        return
    if args.workload == "statics-q4":
        bench_statics(args)
        return
    if args.workload == "periodic-j2":
        bench_periodic_j2(args)
        return
    if args.workload == "quad-api":
        import goosefem_amd as gf

        ceiling, probes = copy_ceiling(gf)
        print(json.dumps({"quad4": api_kernels("quad4", args.nx), "hex8": api_kernels("hex8", args.hex_n),
                          "copy_ceiling_GBs": ceiling, "stream_probes_GBs": probes}))
        return

    nx = ny = args.nx
    S = setup(nx, ny)
    gf, X = S["gf"], S["X"]
    prec = gf.F32 if args.precision == "f32" else gf.F64
    ndof = S["vec"].ndof
    t_all = time.perf_counter()

    # errors of the FP32 and FP64-fast element bodies against the FP64 reference-order path over the
    # workload's 1000 steps (and over --steps when that differs)
    err = {}
    for nst in ([] if args.no_errors else sorted({1000, args.steps})):
        traj = {}
        for p_ in (gf.F64, gf.F32, gf.F64_FAST):
            X.set_state(S["u0"], S["v0"], S["a0"])
            X.step(nst, p_, store_fint=True)
            torch.cuda.synchronize()
            traj[p_] = (X.state()[0].cpu().numpy(), X.fint().cpu().numpy())
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
        u64, f64 = traj[gf.F64]
        err[str(nst)] = {"f32": {"u_rel": rel(traj[gf.F32][0], u64), "fint_rel": rel(traj[gf.F32][1], f64)},
                         "f64_fast": {"u_rel": rel(traj[gf.F64_FAST][0], u64),
                                      "fint_rel": rel(traj[gf.F64_FAST][1], f64)}}
    main_err = err["1000"]["f32" if prec == gf.F32 else "f64_fast"] if err else {"u_rel": None, "fint_rel": None}

    # timed runs: warmup, then exactly `steps` steps, for each precision
    timing = {}
    for p_ in (prec, gf.F64 if prec == gf.F32 else gf.F32, gf.F64_FAST):
        X.set_state(S["u0"], S["v0"], S["a0"])
        X.step(args.warmup, p_)
        torch.cuda.synchronize()
        timing[p_] = time_steps(X, args.steps, p_, gf)
    wall, ktime = timing[prec]

    B = canonical_bytes_per_step(nx, ny)
    per_launch = ktime / args.steps
    achieved = B / per_launch / 1e9
    value = ndof * args.steps / wall
    traffic = load_traffic(nx)
    ceiling, ceilings = copy_ceiling(gf)
    name = {gf.F64: "f64", gf.F32: "f32", gf.F64_FAST: "f64_fast"}
    legs = {}
    for p_, (w_, k_) in timing.items():
        if p_ == prec:
            continue
        legs[name[p_]] = {"value": ndof * args.steps / w_, "avg_launch_us": k_ / args.steps * 1e6,
                          "achieved_GBs": B / (k_ / args.steps) / 1e9,
                          "frac": B / (k_ / args.steps) / 1e9 / HBM_PEAK_GBS,
                          "body": {"f64": "reference expression order, bit-exact with the oracle",
                                   "f32": "FP32 element body, FP64 assembly and state",
                                   "f64_fast": "FP64 closed-form rectangle body with FMA (1e-14 of f64)"}[name[p_]]}
    res = {
        "metric": "DOF-updates/s (Quad4 explicit velocity-Verlet)",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 element / f64 assembly+state" if prec == gf.F32 else "f64",
        "data": "synthetic (seeded log-normal G, driven top edge)",
        "config": {"workload": f"Quad4 Regular {nx}x{ny} plane strain, 2x2 Gauss, velocity-Verlet dt=0.01",
                   "ndof": ndof, "nelem": nx * ny, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_vv_structured", "algorithmic_bytes_per_launch": B,
                     "avg_launch_us": per_launch * 1e6,
                     # the bytes the kernel moves (counter traffic, which is below B: Regular coordinates
                     # are generated, a uniform K is a scalar, no connectivity is read) over its time
                     "moved_GBs": traffic / per_launch / 1e9 if traffic else None,
                     "frac_moved": traffic / per_launch / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "copy_ceiling_GBs": ceiling, "stream_probes_GBs": ceilings,
                     "frac_of_copy_ceiling": achieved / ceiling,
                     "frac_moved_of_copy_ceiling": traffic / per_launch / 1e9 / ceiling if traffic else None},
        "other_precision": legs,
        "error_vs_fp64": {"u_rel": main_err["u_rel"], "fint_rel": main_err["fint_rel"], "steps": 1000,
                          "by_steps": err},
    }
    if not args.no_general:
        res["general_mesh"] = explicit_general(nx, ny, args.steps, args.warmup)
    if not args.no_cpu:
        v_cpu, t_cpu = cpu_baseline(S, nx, ny, args.cpu_steps)
        res["cpu_baseline"] = {"value": v_cpu, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
                               "sample": f"{args.cpu_steps} velocity-Verlet steps of the same {nx}x{ny} "
                                         f"workload, FP64 oracle, one thread, {t_cpu:.1f} s"}
        v_par, t_par, nthr = cpu_baseline_parallel(S, args.cpu_steps * 2)
        res["cpu_baseline_all_cores"] = {"value": v_par, "unit": "DOF-updates/s", "cores": nthr, "kind": "port",
                                         "host_cpus": os.cpu_count(),
                                         "sample": f"{args.cpu_steps * 2} velocity-Verlet steps of the same "
                                                   f"{nx}x{ny} workload, FP64 oracle on {nthr} OpenMP threads "
                                                   f"(bit-identical to the one-thread run), {t_par:.1f} s"}
    del S, X
    torch.cuda.empty_cache()
    if not args.no_blocks:
        # the metric's other workloads: learned vs plain PCG (all held-out fields) and the Hex8 stiffness
        if os.path.exists(args.ws_weights):
            pl = pcg_learned(args.ws_weights, args.ws_tests, cpu=not args.no_cpu, cpu_seconds=5.0)
            res["pcg_learned"] = {k: pl[k] for k in ("plain_s", "learned_s", "speedup_vs_plain", "iterations",
                                                     "inference_ms", "max_true_relres", "config", "per_field")}
            res["pcg_learned"]["unit"] = "s to true relative residual 1e-10 (inference included)"
            if "cpu_baseline" in pl:
                res["pcg_learned"]["cpu_baseline"] = pl["cpu_baseline"]
            torch.cuda.empty_cache()
        res["api_kernels"] = {"quad4": api_kernels("quad4", nx), "hex8": api_kernels("hex8", args.hex_n)}
        for blk in res["api_kernels"].values():  # the same measured ceiling as the headline's
            for row in (blk["f64"], blk["f32"]):
                for r in row.values():
                    r["frac_of_copy_ceiling"] = r["achieved_GBs"] / ceiling
        hx = hex8_stiffness(args.hex_n, cpu=not args.no_cpu, csr=False, cpu_seconds=5.0)
        res["hex8_stiffness"] = {k: hx[k] for k in ("value", "unit", "config", "roofline", "mfma_busy_est",
                                                    "fp64_stiffness", "fp32_io", "tangent_ms", "error_vs_fp64")}
        res["hex8_stiffness"]["roofline"]["frac_of_copy_ceiling"] = hx["roofline"]["achieved"] / ceiling
        if "cpu_baseline" in hx:
            res["hex8_stiffness"]["cpu_baseline"] = hx["cpu_baseline"]
    res["bench_wall_s"] = time.perf_counter() - t_all
    print(json.dumps(res))


def copy_ceiling(gf, gib=4):
    """Measured HBM ceilings (GB/s of bytes read + written) of the library's streaming probes over
    4 GiB arrays: copy (best of plain and non-temporal), add (2 reads : 1 write) and read-only.
    Returns (ceiling, all): the ceiling is the faster of copy and add, the two mixes that bracket the
    explicit step's 1.4 : 1 read : write traffic."""
    import torch

    n = (gib << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    rate = {}
    for mode, name, nbytes in ((0, "copy", 16 * n), (1, "copy_nt", 16 * n), (2, "add", 24 * n), (3, "read", 8 * n)):
        t = _timed(lambda: gf._call("gf_stream_probe", mode, gf._dp(a), gf._dp(b), gf._dp(c), n, gf._stream()), 10)
        rate[name] = nbytes / t / 1e9
    del a, b, c
    torch.cuda.empty_cache()
    rate["copy"] = max(rate["copy"], rate.pop("copy_nt"))
    return max(rate["copy"], rate["add"]), rate


def explicit_general(nx, ny, steps, warmup):
    """The general-mesh explicit path (one kernel per step: per spatially compact block of 256
    nodes, the forces of the elements touching them, geometry recomputed from the coordinates, into
    LDS, then each node's ordered sum and update) on (1) the same Regular mesh with its elements
    renumbered by a random permutation (which defeats the structured-grid detection) and (2) a
    FineLayer mesh of about the same size (3.95 M elements, coarsening layers above and below 1800
    fine layers).  Algorithmic bytes per step: 64 ndof + 20 nelem, as for the headline."""
    import torch

    import goosefem_amd as gf

    out = {}
    meshes = (("renumbered", gf.Mesh.Quad4.Regular(nx, ny), True),
              ("finelayer", gf.Mesh.Quad4.FineLayer(2187, 2400, 1.0, 1800), False))
    for key, mesh, shuffle in meshes:
        coor, dofs = mesh.coor, mesh.dofs
        conn = mesh.conn[np.random.default_rng(5).permutation(mesh.nelem)] if shuffle else mesh.conn
        vec = gf.Vector(conn, dofs)
        inv = np.ones(vec.ndof)
        G = np.exp(0.3 * gf.randn(2, mesh.nelem)).astype(np.float32)
        K = np.full(mesh.nelem, 10.0, dtype=np.float32)
        X = gf.ExplicitQuad4(vec, coor, K, G, inv, 0.01)
        assert not X.structured
        u0 = 1e-3 * gf.randn(1, vec.ndof)
        B = 64 * vec.ndof + 20 * mesh.nelem
        row = {"nelem": mesh.nelem, "ndof": vec.ndof, "algorithmic_bytes_per_step": B, "path": X.path}
        for p_, nm in ((gf.F32, "f32"), (gf.F64, "f64"), (gf.F64_FAST, "f64_fast")):
            X.set_state(u0, np.zeros(vec.ndof), np.zeros(vec.ndof))
            X.step(warmup, p_)
            wall, kt = time_steps(X, steps, p_, gf)
            row[nm] = {"ms_per_step": kt / steps * 1e3, "DOF_updates_per_s": vec.ndof * steps / wall,
                       "achieved_GBs": B / (kt / steps) / 1e9, "frac": B / (kt / steps) / 1e9 / HBM_PEAK_GBS}
        out[key] = row
        del X, vec
        torch.cuda.empty_cache()
    out["workload"] = (f"general path: Quad4 Regular {nx}x{ny} with elements renumbered (random permutation); "
                       "Quad4 FineLayer(2187, 2400, nfine 1800)")
    return out


def api_kernels(kind, n):
    """The unfused quadrature calls of the reference's API (SymGradN_vector, GradN_vector,
    Int_gradN_dot_tensor2_dV; Element.h:634-717, 807-825) on Quad4 planar n x n or Hex8 n^3, FP64
    and FP32 element data, each timed with HIP events.  Algorithmic bytes per element (FP64; half
    for FP32): Quad4 gradient dN/dx 256 + elemvec 64 read, qtensor 288 written; integral dN/dx 256 +
    dV 32 + qtensor 288 read, elemvec 64 written.  Hex8: 1536 + 192 / 576 and 1536 + 64 + 576 / 192."""
    import torch

    import goosefem_amd as gf

    mesh = gf.Mesh.Hex8.Regular(n, n, n) if kind == "hex8" else gf.Mesh.Quad4.Regular(n, n)
    vec = gf.Vector(mesh.conn, mesh.dofs)
    x = vec.AsElement(mesh.coor)
    quad = gf.Element.Hex8.Quadrature(x) if kind == "hex8" else gf.Element.Quad4.QuadraturePlanar(x)
    del x
    E, nip, nne, nd = mesh.nelem, quad.nip, quad.nne, quad.ndim
    bytes_grad = nip * nne * nd * 8 + nne * nd * 8 + nip * 9 * 8
    bytes_int = nip * nne * nd * 8 + nip * 8 + nip * 9 * 8 + nne * nd * 8
    out = {"workload": f"{'Hex8 ' + str(n) + '^3' if kind == 'hex8' else 'Quad4 planar ' + str(n) + '^2'}",
           "nelem": E, "bytes_per_element_f64": {"gradient": bytes_grad, "integral": bytes_int}}
    u = torch.from_numpy(gf.randn(11, E * nne * nd)).cuda().reshape(E, nne, nd)
    for dt, nm, scale in ((torch.float64, "f64", 1.0), (torch.float32, "f32", 0.5)):
        ue = u.to(dt)
        eps = quad.SymGradN_vector(ue)
        fe = quad.Int_gradN_dot_tensor2_dV(eps)
        t_sym = _timed(lambda: quad.SymGradN_vector(ue, out=eps), 10)
        t_grad = _timed(lambda: quad.GradN_vector(ue, out=eps), 10)
        t_int = _timed(lambda: quad.Int_gradN_dot_tensor2_dV(eps, out=fe), 10)
        row = {}
        for key, t, b in (("SymGradN_vector", t_sym, bytes_grad), ("GradN_vector", t_grad, bytes_grad),
                          ("Int_gradN_dot_tensor2_dV", t_int, bytes_int)):
            gbs = scale * b * E / t / 1e9
            row[key] = {"us": t * 1e6, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
            if kind == "hex8" and nm == "f64" and key != "Int_gradN_dot_tensor2_dV":
                # the FP64 Hex8 gradients recompute dN/dx from the element coordinates: they move
                # coordinates 192 + u_e 192 + qtensor 576 B per element, not the stored 1536 B of dN/dx
                moved = nne * nd * 8 * 2 + nip * 9 * 8
                row[key].update(geometry="recomputed from the element coordinates",
                                moved_bytes_per_element=moved, frac_moved=moved * E / t / 1e9 / HBM_PEAK_GBS)
        out[nm] = row
        del eps, fe, ue
    del u, quad, vec
    if kind == "hex8":
        # the per-point isotropic tangent (81 entries per point) that feeds the stiffness
        Kq = torch.full((E, nip), 10.0, dtype=torch.float64, device="cuda")
        Gq = 0.5 + torch.rand((E, nip), dtype=torch.float64, device="cuda")
        for dt, nm, sz in ((None, "f64", 8), (torch.float32, "f32", 4)):
            C = gf.elastic_tangent(Kq, Gq, dtype=dt)
            t = _timed(lambda: gf._call("gf_material_elastic_tangent_f32" if sz == 4 else "gf_material_elastic_tangent",
                                        E * nip, 3, gf._dp(Kq), gf._dp(Gq), gf._dp(C), gf._stream()), 10)
            b = E * nip * (81 * sz + 16)  # writes C, reads K and G
            out[nm]["elastic_tangent"] = {"us": t * 1e6, "achieved_GBs": b / t / 1e9, "frac": b / t / 1e9 / HBM_PEAK_GBS}
            del C
        del Kq, Gq
    torch.cuda.empty_cache()
    return out


def cpu_baseline_parallel(S, nsteps):
    """The oracle's FP64 restatement on all host threads (OpenMP), same workload."""
    import oracle

    ov = oracle.Vector(S["conn"], S["dofs"])
    q = oracle.Quadrature(oracle.QUAD4_PLANAR, ov.AsElement(S["coor"]))
    u, v, a = S["u0"].copy(), S["v0"].copy(), S["a0"].copy()
    t0 = time.perf_counter()
    _, nthr = oracle.vv_steps_parallel(nsteps, oracle.QUAD4_PLANAR, S["conn"], S["dofs"], q.dNx, q.dV(),
                                       S["K"].astype(np.float64), S["G"].astype(np.float64), S["inv"], 0.01, u, v, a)
    dt = time.perf_counter() - t0
    return ov.ndof * nsteps / dt, dt, nthr


if __name__ == "__main__":
    main()
