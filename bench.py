"""Benchmark: Quad4 explicit dynamics on Mesh::Quad4::Regular(2048, 2048), plane strain, 2x2 Gauss,
per-element log-normal shear modulus (sigma 0.3), K = 10, lumped mass (density 1), bottom edge
fixed, top edge driven at x-velocity 1e-4, dt = 0.01.  One step = one velocity-Verlet step over the
whole mesh (fused kernel).  Metric: DOF-updates/s = ndof * steps / wall time.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--nx 2048] [--no-cpu]

Prints one JSON line.  Inputs are synthetic (SplitMix64/Box-Muller generator, seed 2), resident
in HBM before the timed region.
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


def load_traffic(nx):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if it matches this size."""
    p = os.path.join(ROOT, "profiles", "explicit_vv_quad4_2048_traffic.json")
    try:
        d = json.load(open(p))
        if d.get("nx") == nx:
            return d.get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (PMC counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the path shards by mesh slabs "
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

    nx = ny = args.nx
    S = setup(nx, ny)
    gf, X = S["gf"], S["X"]
    prec = gf.F32 if args.precision == "f32" else gf.F64
    ndof = S["vec"].ndof

    # FP64 reference trajectory of the same length (error of the FP32 element kernels vs FP64)
    X.step(args.steps, gf.F64, store_fint=True)
    torch.cuda.synchronize()
    u64 = X.state()[0].cpu().numpy()
    f64 = X.fint().cpu().numpy()
    X.set_state(S["u0"], S["v0"], S["a0"])
    X.step(args.steps, gf.F32, store_fint=True)
    torch.cuda.synchronize()
    u32 = X.state()[0].cpu().numpy()
    f32 = X.fint().cpu().numpy()
    err_u = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    err_f = float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))

    # timed run: warmup, then exactly `steps` steps
    X.set_state(S["u0"], S["v0"], S["a0"])
    X.step(args.warmup, prec)
    torch.cuda.synchronize()
    wall, ktime = time_steps(X, args.steps, prec, gf)
    other = gf.F64 if prec == gf.F32 else gf.F32
    X.set_state(S["u0"], S["v0"], S["a0"])
    X.step(args.warmup, other)
    wall_o, ktime_o = time_steps(X, args.steps, other, gf)

    B = canonical_bytes_per_step(nx, ny)
    per_launch = ktime / args.steps
    achieved = B / per_launch / 1e9
    value = ndof * args.steps / wall
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(nx),
                     "kernel": "k_vv_structured", "algorithmic_bytes_per_launch": B,
                     "avg_launch_us": per_launch * 1e6},
        "other_precision": {"dtype": "f64" if prec == gf.F32 else "f32",
                            "value": ndof * args.steps / wall_o,
                            "avg_launch_us": ktime_o / args.steps * 1e6,
                            "achieved_GBs": B / (ktime_o / args.steps) / 1e9},
        "error_vs_fp64": {"u_rel": err_u, "fint_rel": err_f, "steps": args.steps},
    }
    if not args.no_cpu:
        v_cpu, t_cpu = cpu_baseline(S, nx, ny, args.cpu_steps)
        res["cpu_baseline"] = {"value": v_cpu, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
                               "sample": f"{args.cpu_steps} velocity-Verlet steps of the same {nx}x{ny} "
                                         f"workload, FP64 oracle, one thread, {t_cpu:.1f} s"}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
