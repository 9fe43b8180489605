"""Learned warm start: the BF16 MFMA convolution, pooling, loss and Adam kernels against the FP64
CPU restatement (oracle/convnet.py), the whole network's forward and gradients, determinism, and
the warm-started PCG (same true-residual tolerance as from a zero guess)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL_F32 = 2e-5      # FP32 outputs of BF16 x BF16 products with FP32 accumulation, vs FP64
ULP_BF16 = 2.0 ** -8  # one bfloat16 rounding of an output
TOL_NET = 5e-2      # whole network: BF16 activations and gradients at every layer


@pytest.fixture(scope="module")
def gf():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    import goosefem_amd

    goosefem_amd.lib()
    return goosefem_amd


@pytest.fixture(scope="module")
def ws(gf):
    from goosefem_amd import warmstart

    return warmstart


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def bf16_rand(shape, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return (scale * torch.randn(*shape, dtype=torch.float64, generator=g)).to(torch.bfloat16)


def cpu64(t):
    return t.detach().cpu().to(torch.float64)


def packed(ws, w, cout, cin, cin_s, cout_s):
    """master FP32 [9][cout][cin] -> (wf, wb, coutp, cinp) on the device"""
    coutp, cinp = ws.coutp_of(cout_s), ws.coutp_of(cin_s)
    wf = torch.zeros(9 * coutp * cin_s, dtype=torch.bfloat16, device="cuda")
    wb = torch.zeros(9 * cinp * cout_s, dtype=torch.bfloat16, device="cuda")
    wd = w.to(torch.float32).cuda().contiguous()
    ws._call("gf_conv_pack", ws._i32(cout), ws._i32(cin), ws._dp(wd), ws._i32(coutp), ws._i32(cin_s), ws._dp(wf),
             ws._i32(cinp), ws._i32(cout_s), ws._dp(wb), ws._stream())
    return wf, wb, coutp, cinp


def call_conv(ws, x1, c1, up1, x2, c2, N, H, W, wpk, coutp, bias, cout, flags=0, y=None, ysplit=None, y2=None,
              y32=None, mask=None):
    ws._call("gf_conv3x3", ws._dp(x1), ws._i32(c1), ws._i32(up1), ws._dp(x2), ws._i32(c2), ws._i64(N * H * W),
             ws._i32(H), ws._i32(W), ws._dp(wpk), ws._i32(coutp), ws._dp(bias), ws._i32(cout), ws._i32(flags),
             ws._dp(y), ws._i32(ysplit if ysplit is not None else cout), ws._dp(y2), ws._dp(y32), ws._dp(mask),
             ws._stream())


CONV_CASES = [
    # N, H, W, c1, c2, up1, cout
    (1, 16, 16, 16, 0, 0, 16),
    (2, 8, 16, 32, 16, 1, 32),
    (1, 32, 32, 16, 0, 0, 2),
    (1, 8, 8, 128, 128, 1, 128),
    (1, 16, 16, 64, 0, 0, 96),
    (3, 4, 8, 48, 0, 0, 200),
]


@pytest.mark.parametrize("N,H,W,c1,c2,up1,cout", CONV_CASES)
def test_conv_forward_fp32_output(ws, N, H, W, c1, c2, up1, cout):
    from oracle import convnet

    x1 = bf16_rand((N, H >> up1, W >> up1, c1), 1)
    x2 = bf16_rand((N, H, W, c2), 2) if c2 else None
    cin = c1 + c2
    w = (torch.randn(9, cout, cin, dtype=torch.float64, generator=torch.Generator().manual_seed(3)) /
         np.sqrt(9 * cin)).to(torch.float32)
    b = torch.linspace(-0.5, 0.5, cout, dtype=torch.float32)
    wf, _, coutp, _ = packed(ws, w, cout, cin, cin, cout)
    y32 = torch.zeros(N * H * W * cout, dtype=torch.float32, device="cuda")
    call_conv(ws, x1.cuda(), c1, up1, x2.cuda() if c2 else None, c2, N, H, W, wf, coutp, b.cuda(), cout, y32=y32)
    xin = cpu64(x1)
    if up1:
        xin = convnet.upsample2(xin)
    if c2:
        xin = torch.cat([xin, cpu64(x2)], dim=3)
    ref = convnet.conv3x3(xin, convnet.bf16(w.double()), b.double())
    assert rel(y32.cpu().reshape(N, H, W, cout), ref) < TOL_F32


def test_conv_bf16_epilogue_relu_mask_accum_split(ws):
    from oracle import convnet

    N, H, W, c1, cout, split = 2, 16, 16, 32, 48, 16
    x1 = bf16_rand((N, H, W, c1), 4)
    w = (torch.randn(9, cout, c1, dtype=torch.float64, generator=torch.Generator().manual_seed(5)) / 12).float()
    b = torch.linspace(-0.2, 0.2, cout, dtype=torch.float32)
    wf, _, coutp, _ = packed(ws, w, cout, c1, c1, cout)
    mask = bf16_rand((N, H, W, cout), 6)
    old = bf16_rand((N, H, W, cout), 7)
    ya = old[..., :split].contiguous().cuda()
    yb = old[..., split:].contiguous().cuda()
    call_conv(ws, x1.cuda(), c1, 0, None, 0, N, H, W, wf, coutp, b.cuda(), cout, flags=ws.CF_RELU | ws.CF_ACCUM,
              y=ya, ysplit=split, y2=yb, mask=mask.cuda())
    ref = convnet.conv3x3(cpu64(x1), convnet.bf16(w.double()), b.double()) + cpu64(old)
    ref = torch.relu(ref) * (cpu64(mask) > 0)
    got = torch.cat([cpu64(ya), cpu64(yb)], dim=3)
    err = (got - ref).abs()
    assert bool((err <= 2 * ULP_BF16 * ref.abs() + 1e-6).all()), float(err.max())


@pytest.mark.parametrize("N,H,W,c1,c2,up1,cout", CONV_CASES[:5])
def test_conv_data_and_weight_gradients(ws, N, H, W, c1, c2, up1, cout):
    """data gradient (tap-flipped transposed weights through the same kernel, then the 2x2 block
    sum for the upsampled input) and weight/bias gradients vs autograd of the FP64 restatement"""
    from oracle import convnet

    cin = c1 + c2
    cout_s = max(16, (cout + 15) // 16 * 16)
    x1 = bf16_rand((N, H >> up1, W >> up1, c1), 11)
    x2 = bf16_rand((N, H, W, c2), 12) if c2 else None
    w = (torch.randn(9, cout, cin, dtype=torch.float64, generator=torch.Generator().manual_seed(13)) /
         np.sqrt(9 * cin)).float()
    gy = torch.zeros(N, H, W, cout_s, dtype=torch.bfloat16)
    gy[..., :cout] = bf16_rand((N, H, W, cout), 14)
    wf, wb, coutp, cinp = packed(ws, w, cout, cin, cin, cout_s)
    # data gradient
    gfull = torch.zeros(N * H * W * cin, dtype=torch.float32, device="cuda")
    call_conv(ws, gy.cuda(), cout_s, 0, None, 0, N, H, W, wb, cinp, None, cin, y32=gfull)
    # autograd reference
    a1 = cpu64(x1).requires_grad_()
    xin = convnet.upsample2(a1) if up1 else a1
    a2 = None
    if c2:
        a2 = cpu64(x2).requires_grad_()
        xin = torch.cat([xin, a2], dim=3)
    wr = convnet.bf16(w.double()).requires_grad_()
    br = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    out = convnet.conv3x3(xin, wr, br)
    out.backward(cpu64(gy)[..., :cout])
    gfull = gfull.cpu().double().reshape(N, H, W, cin)
    if up1:
        gup = gfull[..., :c1]
        g1 = gup.reshape(N, H // 2, 2, W // 2, 2, c1).sum(dim=(2, 4))
    else:
        g1 = gfull[..., :c1]
    assert rel(g1, a1.grad) < TOL_F32
    if c2:
        assert rel(gfull[..., c1:], a2.grad) < TOL_F32
    # weight and bias gradient
    ns, nf = ws.ctypes.c_int(0), ws.ctypes.c_int64(0)
    ws._call("gf_conv3x3_wgrad_work", ws._i64(N * H * W), ws._i32(cout_s), ws._i32(cin), ws.ctypes.byref(ns),
             ws.ctypes.byref(nf))
    work = torch.zeros(nf.value, dtype=torch.float32, device="cuda")
    dw = torch.zeros(9 * cout * cin, dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    # device copies held in locals: a temporary freed inside the argument list would hand its
    # memory to the next allocation before the kernel reads it
    x1d, gyd = x1.cuda(), gy.cuda()
    x2d = x2.cuda() if c2 else None
    ws._call("gf_conv3x3_wgrad", ws._dp(x1d), ws._i32(c1), ws._i32(up1), ws._dp(x2d),
             ws._i32(c2), ws._i64(N * H * W), ws._i32(H), ws._i32(W), ws._dp(gyd), ws._i32(cout_s),
             ws._i32(cout), ws._i32(cin), ws._dp(work), ws._i64(work.numel()), ws._dp(dw), ws._dp(db), ws._stream())
    assert rel(dw.cpu().reshape(9, cout, cin), wr.grad) < TOL_F32
    assert rel(db.cpu(), br.grad) < TOL_F32


def test_pool_kernels(ws):
    from oracle import convnet

    N, H, W, C = 2, 8, 16, 32
    x = bf16_rand((N, H, W, C), 21)
    xd = x.cuda()
    out = torch.zeros(N * H * W * C // 4, dtype=torch.bfloat16, device="cuda")
    ws._call("gf_avgpool2", ws._dp(xd), ws._i64(N), ws._i32(H), ws._i32(W), ws._i32(C), ws._dp(out),
             ws._stream())
    ref = convnet.avgpool2(cpu64(x))
    assert bool(((cpu64(out).reshape(ref.shape) - ref).abs() <= ULP_BF16 * ref.abs() + 1e-7).all())
    # unpool: din = old + 0.25 dout (fine) masked by act
    dout = bf16_rand((N, H // 2, W // 2, C), 22)
    old = bf16_rand((N, H, W, C), 23)
    act = bf16_rand((N, H, W, C), 24)
    din = old.clone().cuda()
    doutd, actd = dout.cuda(), act.cuda()
    ws._call("gf_unpool2", ws._dp(doutd), ws._i64(N), ws._i32(H), ws._i32(W), ws._i32(C), ws.ctypes.c_float(0.25),
             ws._i32(1), ws._dp(actd), ws._dp(din), ws._stream())
    ref = (cpu64(old) + 0.25 * convnet.upsample2(cpu64(dout))) * (cpu64(act) > 0)
    assert bool(((cpu64(din) - ref).abs() <= ULP_BF16 * ref.abs() + 1e-7).all())
    # sumpool: coarse = 2x2 block sums masked by coarse act
    cact = bf16_rand((N, H // 2, W // 2, C), 25)
    so = torch.zeros(N * H * W * C // 4, dtype=torch.bfloat16, device="cuda")
    cactd = cact.cuda()
    ws._call("gf_sumpool2", ws._dp(xd), ws._i64(N), ws._i32(H), ws._i32(W), ws._i32(C), ws._dp(cactd),
             ws._dp(so), ws._stream())
    ref = 4 * convnet.avgpool2(cpu64(x)) * (cpu64(cact) > 0)
    assert bool(((cpu64(so).reshape(ref.shape) - ref).abs() <= ULP_BF16 * ref.abs() + 1e-7).all())


def test_adam_closed_form(ws):
    n = 1000
    rng = np.random.default_rng(0)
    w0 = rng.standard_normal(n).astype(np.float32)
    g1, g2 = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    w = torch.from_numpy(w0.copy()).cuda()
    m = torch.zeros(n, dtype=torch.float32, device="cuda")
    v = torch.zeros_like(m)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    for step, g in ((1, g1), (2, g2)):
        gd = torch.from_numpy(g).cuda()
        ws._call("gf_adam", ws._i64(n), ws._dp(w), ws._dp(gd), ws._dp(m), ws._dp(v),
                 ws.ctypes.c_float(lr), ws.ctypes.c_float(b1), ws.ctypes.c_float(b2), ws.ctypes.c_float(eps),
                 ws._i64(step), ws._stream())
    M, V, Wr = np.zeros(n), np.zeros(n), w0.astype(np.float64)
    for step, g in ((1, g1), (2, g2)):
        M = b1 * M + (1 - b1) * g
        V = b2 * V + (1 - b2) * g.astype(np.float64) ** 2
        Wr = Wr - lr * (M / (1 - b1 ** step)) / (np.sqrt(V / (1 - b2 ** step)) + eps)
    assert np.allclose(w.cpu().numpy(), Wr, rtol=1e-5, atol=1e-6)


def net_and_params(ws, channels, seed=3):
    net = ws.WarmStartNet(channels, seed=seed)
    params = []
    for c in net.layers():
        wq = c.w.cpu().double().reshape(9, c.cout, c.cin)
        params.append((convnet_bf16(wq).requires_grad_(), c.b.cpu().double().requires_grad_()))
    return net, params


def convnet_bf16(t):
    from oracle import convnet

    return convnet.bf16(t)


@pytest.mark.parametrize("channels,B,H,W", [((16, 32), 2, 32, 32), ((16, 32, 64), 1, 32, 64)])
def test_network_forward_and_gradients_vs_oracle(ws, channels, B, H, W):
    from oracle import convnet

    net, params = net_and_params(ws, channels)
    # non-zero biases so that every bias path is exercised
    for i, c in enumerate(net.layers()):
        c.b.copy_(torch.linspace(-0.05, 0.05, c.cout, device="cuda"))
        params[i] = (params[i][0], c.b.cpu().double().requires_grad_())
    G = [torch.from_numpy(np.exp(0.4 * np.random.default_rng(b).standard_normal(H * W))).cuda() for b in range(B)]
    sigma = 0.4
    net.set_input(B, H, W, G, 0.0, 1.0 / sigma)
    pred = net.forward(B, H, W)
    z = torch.stack([torch.log(g.cpu()) / sigma for g in G]).reshape(B, H, W, 1)
    zq = convnet.bf16(z)
    ref = convnet.unet_forward(zq, params, channels, round_bf16=True)
    got = pred.cpu().double().reshape(B, H, W, 2)
    assert rel(got, ref.detach()) < TOL_NET
    tgt = torch.randn(B, H, W, 2, dtype=torch.float64, generator=torch.Generator().manual_seed(9))
    loss = net.loss_grad(B, H, W, tgt.float().cuda().reshape(-1))
    net.backward(B, H, W)
    lref = ((ref.detach() - tgt) ** 2).sum()
    assert abs(loss.item() - ((got - tgt) ** 2).sum().item()) < 1e-4 * lref.item()
    (((ref - tgt) ** 2).mean()).backward()
    for i, ((wr, br), c) in enumerate(zip(params, net.layers())):
        e_w = rel(c.dw.cpu().reshape(9, c.cout, c.cin), wr.grad)
        e_b = rel(c.db.cpu(), br.grad)
        assert e_w < TOL_NET and e_b < TOL_NET, (i, e_w, e_b)


def test_training_step_deterministic_and_loss_decreases(ws):
    B, H, W = 2, 32, 32
    out = []
    for rep in range(2):
        net = ws.WarmStartNet((16, 32), seed=5)
        G = [torch.from_numpy(np.exp(0.4 * np.random.default_rng(b).standard_normal(H * W))).cuda() for b in range(B)]
        tgt = torch.from_numpy(np.random.default_rng(7).standard_normal(B * H * W * 2).astype(np.float32)).cuda()
        net.set_input(B, H, W, G, 0.0, 2.5)
        losses = []
        for it in range(30):
            net.forward(B, H, W)
            losses.append(net.loss_grad(B, H, W, tgt).item())
            net.backward(B, H, W)
            net.adam(3e-3)
        out.append((losses, net.layers()[0].w.cpu().numpy().copy()))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
    assert out[0][0][-1] < out[0][0][0]


def test_warm_start_pcg_same_tolerance(gf, ws):
    """Round trip solution -> centred grid -> initial guess is exact up to FP32 rounding, the
    warm-started solve from it converges in a few iterations, and an untrained network's guess still
    converges to the same true residual."""
    n = 32
    mesh = gf.Mesh.Quad4.Regular(n, n)
    ctrl = gf.Tyings.Control(mesh.coor, mesh.dofs)
    iip = np.concatenate((ctrl.controlDofs.ravel(), ctrl.dofs[mesh.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, mesh.nodesPeriodic, iip)
    vec = gf.Vector(mesh.conn, ty.dofs)
    quad = gf.Element.Quad4.QuadraturePlanar(gf.Vector(mesh.conn, mesh.dofs).AsElement(mesh.coor))
    G = np.exp(0.4 * gf.randn(77, n * n))
    K = np.full(n * n, 10.0)
    P = gf.TyingsPCG(vec, quad, K, G, ty)
    net = ws.WarmStartNet((16, 32), seed=1)
    W = ws.PeriodicShearWarmStart(P, ty, n, net, target_scale=0.1)
    gamma = 1e-3
    b = W.rhs(gamma)
    x, it0, res0 = P.solve(b, tol=1e-10)
    grid = torch.zeros(n * n * 2, dtype=torch.float32, device="cuda")
    W.target(x, gamma, grid)
    x0 = torch.zeros(ty.nnu, dtype=torch.float64, device="cuda")
    gf._call("gf_grid_to_dofs", gf._i64(n * n * 2), gf._i32(2), gf._dp(grid), gf._dp(W.idx), gf._dp(W.affine),
             gf._f64(W.target_scale), gf._f64(gamma), gf._dp(x0), gf._stream())
    assert rel(x0.cpu(), x.cpu()) < 1e-5
    _, it1, res1 = P.solve(b, x0=x0, tol=1e-10)
    assert res1 <= 1e-9 and it1 < 0.7 * it0
    xg = W.predict(torch.from_numpy(G).cuda(), gamma)
    _, it2, res2 = P.solve(b, x0=xg, tol=1e-10)
    assert res2 <= 1e-9 and it2 > 0


def test_committed_weights_cut_pcg_iterations_at_512(gf, ws):
    """The committed 512^2 network on a modulus field outside its training seeds: the warm-started
    PCG reaches the same true residual with at least 35% fewer iterations than from a zero guess,
    and at least 20% fewer than from the affine part of the start alone."""
    import os

    path = os.path.join(os.path.dirname(gf.__file__), "data", "warmstart_quad4_512.npz")
    net = ws.WarmStartNet.load(path)
    n = int(net.meta["n"])
    pb = ws.periodic_problem(n, sigma=float(net.meta["sigma"]))
    W = ws.PeriodicShearWarmStart(pb.P, pb.ty, n, net, sigma=pb.sigma)
    G = gf.to_device(np.exp(pb.sigma * gf.randn(123456, n * n)))
    pb.P.set_moduli(gf.to_device(pb.K), G)
    b = W.rhs(1e-3)
    _, it0, r0 = pb.P.solve(b, tol=1e-10)
    x0 = W.predict(G, 1e-3)
    _, it1, r1 = pb.P.solve(b, x0=x0, tol=1e-10)
    assert r0 <= 1e-9 and r1 <= 1e-9
    assert pb.P.start_used == gf.GF_START_GIVEN  # the prediction is kept, not replaced by the fallback
    assert it1 < 0.65 * it0, (it0, it1)  # measured 3277 / 5871 = 0.56
    # what the network adds beyond the affine part of the start (which alone removes about 20%)
    _, it2, _ = pb.P.solve(b, x0=_affine_start(gf, W, 2, 1e-3), tol=1e-10)
    print(f"\n512^2 field 123456: plain {it0}, affine start {it2}, learned start {it1} iterations")
    assert it1 < 0.8 * it2, (it2, it1)  # measured 3277 / 4721 = 0.69


def _affine_start(gf, W, dim, gamma):
    """The start the network's prediction is added to: the affine part alone (a zero fluctuation)."""
    grid = torch.zeros(W.nv * dim, dtype=torch.float32, device="cuda")
    x = torch.zeros(W.ty.nnu, dtype=torch.float64, device="cuda")
    gf._call("gf_grid_to_dofs", gf._i64(W.nv * dim), gf._i32(dim), gf._dp(grid), gf._dp(W.idx), gf._dp(W.affine),
             gf._f64(1.0), gf._f64(gamma), gf._dp(x), gf._stream())
    return x


def test_committed_weights_3d_beat_affine_start(gf, ws):
    """The committed Hex8 96^3 network (stiff spheres, pure shear) on a field outside its training
    seeds: fewer PCG iterations than the affine part of the start alone and than a zero guess, to the
    same true residual."""
    import os

    path = os.path.join(os.path.dirname(gf.__file__), "data", "warmstart_hex8_96.npz")
    net = ws.WarmStartNet.load(path)
    n, dim = int(net.meta["n"]), int(net.meta.get("dim", 3))
    assert dim == 3
    pb = ws.periodic_problem(n, sigma=float(net.meta["sigma"]), dim=3, spheres=True)
    W = ws.PeriodicWarmStart(pb.P, pb.ty, n, net, sigma=pb.sigma, dim=3, shear=str(net.meta.get("shear", "pure")))
    G = gf.to_device(ws.sphere_field(n, 900001, pb.sigma))
    pb.P.set_moduli(gf.to_device(pb.K), G)
    gamma = 5e-4
    b = W.rhs(gamma)
    _, it0, r0 = pb.P.solve(b, tol=1e-10)
    _, it1, r1 = pb.P.solve(b, x0=W.predict(G, gamma), tol=1e-10)
    _, it2, r2 = pb.P.solve(b, x0=_affine_start(gf, W, 3, gamma), tol=1e-10)
    assert max(r0, r1, r2) <= 1e-9
    assert it1 < 0.85 * it2 and it1 < 0.8 * it0, (it0, it2, it1)


# ---- 3-D (3x3x3 taps, 2x2x2 pooling) ---------------------------------------------------------------

def packed3(ws, w, cout, cin, cin_s, cout_s):
    coutp, cinp = ws.coutp_of(cout_s), ws.coutp_of(cin_s)
    wf = torch.zeros(27 * coutp * cin_s, dtype=torch.bfloat16, device="cuda")
    wb = torch.zeros(27 * cinp * cout_s, dtype=torch.bfloat16, device="cuda")
    wd = w.to(torch.float32).cuda().contiguous()
    ws._call("gf_conv_pack3", ws._i32(cout), ws._i32(cin), ws._dp(wd), ws._i32(coutp), ws._i32(cin_s), ws._dp(wf),
             ws._i32(cinp), ws._i32(cout_s), ws._dp(wb), ws._stream())
    return wf, wb, coutp, cinp


CONV3_CASES = [
    # N, Z, H, W, c1, c2, up1, cout
    (1, 4, 6, 8, 16, 0, 0, 16),
    (2, 4, 4, 8, 32, 16, 1, 32),
    (1, 8, 8, 8, 16, 0, 0, 3),
    (1, 4, 4, 4, 64, 64, 1, 64),
]


@pytest.mark.parametrize("N,Z,H,W,c1,c2,up1,cout", CONV3_CASES)
def test_conv3d_forward_and_gradients(ws, N, Z, H, W, c1, c2, up1, cout):
    from oracle import convnet

    cin = c1 + c2
    cout_s = max(16, (cout + 15) // 16 * 16)
    x1 = bf16_rand((N, Z >> up1, H >> up1, W >> up1, c1), 31)
    x2 = bf16_rand((N, Z, H, W, c2), 32) if c2 else None
    w = (torch.randn(27, cout, cin, dtype=torch.float64, generator=torch.Generator().manual_seed(33)) /
         np.sqrt(27 * cin)).float()
    b = torch.linspace(-0.3, 0.3, cout, dtype=torch.float32)
    wf, wb, coutp, cinp = packed3(ws, w, cout, cin, cin, cout_s)
    x1d, bd = x1.cuda(), b.cuda()
    x2d = x2.cuda() if c2 else None
    P = N * Z * H * W
    y32 = torch.zeros(P * cout, dtype=torch.float32, device="cuda")
    ws._call("gf_conv3x3x3", ws._dp(x1d), ws._i32(c1), ws._i32(up1), ws._dp(x2d), ws._i32(c2), ws._i64(P), ws._i32(Z),
             ws._i32(H), ws._i32(W), ws._dp(wf), ws._i32(coutp), ws._dp(bd), ws._i32(cout), ws._i32(0), None,
             ws._i32(cout), None, ws._dp(y32), None, ws._stream())
    a1 = cpu64(x1).requires_grad_()
    xin = convnet.upsample2x2x2(a1) if up1 else a1
    a2 = None
    if c2:
        a2 = cpu64(x2).requires_grad_()
        xin = torch.cat([xin, a2], dim=4)
    wr = convnet.bf16(w.double()).requires_grad_()
    br = b.double().requires_grad_()
    ref = convnet.conv3x3x3(xin, wr, br)
    assert rel(y32.cpu().reshape(N, Z, H, W, cout), ref.detach()) < TOL_F32
    # gradients
    gy = torch.zeros(N, Z, H, W, cout_s, dtype=torch.bfloat16)
    gy[..., :cout] = bf16_rand((N, Z, H, W, cout), 34)
    ref.backward(cpu64(gy)[..., :cout])
    gyd = gy.cuda()
    gfull = torch.zeros(P * cin, dtype=torch.float32, device="cuda")
    ws._call("gf_conv3x3x3", ws._dp(gyd), ws._i32(cout_s), ws._i32(0), None, ws._i32(0), ws._i64(P), ws._i32(Z),
             ws._i32(H), ws._i32(W), ws._dp(wb), ws._i32(cinp), None, ws._i32(cin), ws._i32(0), None, ws._i32(cin),
             None, ws._dp(gfull), None, ws._stream())
    gfull = gfull.cpu().double().reshape(N, Z, H, W, cin)
    g1 = gfull[..., :c1]
    if up1:
        g1 = g1.reshape(N, Z // 2, 2, H // 2, 2, W // 2, 2, c1).sum(dim=(2, 4, 6))
    assert rel(g1, a1.grad) < TOL_F32
    if c2:
        assert rel(gfull[..., c1:], a2.grad) < TOL_F32
    ns, nf = ws.ctypes.c_int(0), ws.ctypes.c_int64(0)
    ws._call("gf_conv3x3x3_wgrad_work", ws._i64(P), ws._i32(cout_s), ws._i32(cin), ws.ctypes.byref(ns),
             ws.ctypes.byref(nf))
    work = torch.zeros(nf.value, dtype=torch.float32, device="cuda")
    dw = torch.zeros(27 * cout * cin, dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    ws._call("gf_conv3x3x3_wgrad", ws._dp(x1d), ws._i32(c1), ws._i32(up1), ws._dp(x2d), ws._i32(c2), ws._i64(P),
             ws._i32(Z), ws._i32(H), ws._i32(W), ws._dp(gyd), ws._i32(cout_s), ws._i32(cout), ws._i32(cin),
             ws._dp(work), ws._i64(work.numel()), ws._dp(dw), ws._dp(db), ws._stream())
    assert rel(dw.cpu().reshape(27, cout, cin), wr.grad) < TOL_F32
    assert rel(db.cpu(), br.grad) < TOL_F32


def test_pool3d_kernels(ws):
    from oracle import convnet

    N, Z, H, W, C = 2, 4, 6, 8, 16
    x = bf16_rand((N, Z, H, W, C), 41)
    xd = x.cuda()
    out = torch.zeros(N * Z * H * W * C // 8, dtype=torch.bfloat16, device="cuda")
    ws._call("gf_avgpool2x2x2", ws._dp(xd), ws._i64(N), ws._i32(Z), ws._i32(H), ws._i32(W), ws._i32(C), ws._dp(out),
             ws._stream())
    ref = convnet.avgpool2x2x2(cpu64(x))
    assert bool(((cpu64(out).reshape(ref.shape) - ref).abs() <= ULP_BF16 * ref.abs() + 1e-7).all())
    dout = bf16_rand((N, Z // 2, H // 2, W // 2, C), 42)
    old = bf16_rand((N, Z, H, W, C), 43)
    act = bf16_rand((N, Z, H, W, C), 44)
    din, doutd, actd = old.clone().cuda(), dout.cuda(), act.cuda()
    ws._call("gf_unpool2x2x2", ws._dp(doutd), ws._i64(N), ws._i32(Z), ws._i32(H), ws._i32(W), ws._i32(C),
             ws.ctypes.c_float(0.125), ws._i32(1), ws._dp(actd), ws._dp(din), ws._stream())
    ref = (cpu64(old) + 0.125 * convnet.upsample2x2x2(cpu64(dout))) * (cpu64(act) > 0)
    assert bool(((cpu64(din) - ref).abs() <= ULP_BF16 * ref.abs() + 1e-7).all())
    cact = bf16_rand((N, Z // 2, H // 2, W // 2, C), 45)
    cactd = cact.cuda()
    so = torch.zeros(N * Z * H * W * C // 8, dtype=torch.bfloat16, device="cuda")
    ws._call("gf_sumpool2x2x2", ws._dp(xd), ws._i64(N), ws._i32(Z), ws._i32(H), ws._i32(W), ws._i32(C),
             ws._dp(cactd), ws._dp(so), ws._stream())
    ref = 8 * convnet.avgpool2x2x2(cpu64(x)) * (cpu64(cact) > 0)
    assert bool(((cpu64(so).reshape(ref.shape) - ref).abs() <= 2 * ULP_BF16 * ref.abs() + 1e-6).all())


def test_network3d_forward_and_gradients_vs_oracle(ws):
    from oracle import convnet

    channels, B, Z, H, W = (16, 32), 1, 8, 8, 8
    net = ws.WarmStartNet(channels, seed=3, dim=3)
    params = []
    for c in net.layers():
        c.b.copy_(torch.linspace(-0.05, 0.05, c.cout, device="cuda"))
        wq = c.w.cpu().double().reshape(27, c.cout, c.cin)
        params.append((convnet.bf16(wq).requires_grad_(), c.b.cpu().double().requires_grad_()))
    G = [torch.from_numpy(np.exp(0.4 * np.random.default_rng(2).standard_normal(Z * H * W))).cuda()]
    net.set_input(B, H, W, G, 0.0, 2.5, Z=Z)
    pred = net.forward(B, H, W, Z=Z)
    z = (torch.log(G[0].cpu()) * 2.5).reshape(1, Z, H, W, 1)
    ref = convnet.unet_forward(convnet.bf16(z), params, channels, round_bf16=True)
    got = pred.cpu().double().reshape(B, Z, H, W, 3)
    assert rel(got, ref.detach()) < TOL_NET
    tgt = torch.randn(B, Z, H, W, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(8))
    net.loss_grad(B, H, W, tgt.float().cuda().reshape(-1), Z=Z)
    net.backward(B, H, W, Z=Z)
    (((ref - tgt) ** 2).mean()).backward()
    for i, ((wr, br), c) in enumerate(zip(params, net.layers())):
        e_w = rel(c.dw.cpu().reshape(27, c.cout, c.cin), wr.grad)
        e_b = rel(c.db.cpu(), br.grad)
        assert e_w < TOL_NET and e_b < TOL_NET, (i, e_w, e_b)


@pytest.mark.parametrize("dim,shape,levels,lam", [(2, (2, 32, 32, 2), 4, 1.0), (2, (1, 16, 48, 2), 3, 4.0),
                                                  (3, (2, 8, 8, 8, 3), 2, 1.0), (3, (1, 4, 8, 16, 3), 2, 0.5)])
def test_mse_multiscale_matches_oracle(ws, dim, shape, levels, lam):
    # gf_mse_multiscale against oracle.convnet.mse_multiscale (itself pinned by closed forms and
    # finite differences): the loss to FP32 summation accuracy, the gradient to one bfloat16 rounding
    from oracle import convnet

    rng = np.random.default_rng(sum(shape) + levels)
    pred = rng.standard_normal(shape).astype(np.float32)
    tgt = (pred + 0.3 * rng.standard_normal(shape) + 0.2).astype(np.float32)  # an offset: mu != 0
    L_ref, g_ref = convnet.mse_multiscale(pred, tgt, lam, levels, dim=dim)
    B, nc, cg = shape[0], shape[-1], 16
    Z, H, W = (1,) + shape[1:3] if dim == 2 else shape[1:4]
    P = B * Z * H * W
    nf = ws._i64()
    ws._call("gf_mse_multiscale_work", ws._i32(dim), ws._i64(B), ws._i32(Z), ws._i32(H), ws._i32(W), ws._i32(nc),
             ws._i32(levels), ws._i32(cg), __import__("ctypes").byref(nf))
    work = torch.zeros(nf.value, dtype=torch.float32, device="cuda")
    g = torch.full((P * cg,), 7.0, dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    pd, td = torch.from_numpy(pred).cuda().reshape(-1), torch.from_numpy(tgt).cuda().reshape(-1)
    scale = 3.0
    ws._call("gf_mse_multiscale", ws._i32(dim), ws._i64(B), ws._i32(Z), ws._i32(H), ws._i32(W), ws._i32(nc),
             ws._i32(levels), ws._f32(lam), ws._dp(pd), ws._dp(td), ws._f32(scale), ws._i32(cg), ws._dp(g),
             ws._dp(work), ws._i64(work.numel()), ws._dp(loss), ws._stream())
    torch.cuda.synchronize()
    assert abs(loss.item() - L_ref) <= 1e-5 * L_ref
    got = cpu64(g).reshape(P, cg)
    want = (scale * g_ref / (P * nc)).reshape(P, nc)
    assert torch.all(got[:, nc:] == 0)
    err = (got[:, :nc] - want).abs()
    assert torch.all(err <= ULP_BF16 * want.abs() + 1e-6 * want.abs().max())
    # a bad level count (2^levels must divide every dimension) is refused
    with pytest.raises(Exception):
        ws._call("gf_mse_multiscale", ws._i32(dim), ws._i64(B), ws._i32(Z), ws._i32(H), ws._i32(W), ws._i32(nc),
                 ws._i32(levels + 4), ws._f32(lam), ws._dp(pd), ws._dp(td), ws._f32(scale), ws._i32(cg), ws._dp(g),
                 ws._dp(work), ws._i64(work.numel()), ws._dp(loss), ws._stream())
