"""Pins the FP64 CPU restatement of the warm-start network (oracle/convnet.py): the convolution
against torch's own circular conv2d (an independent library routine), pooling/upsampling against
closed forms, and the autograd gradients against central finite differences."""

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import convnet


def test_conv3x3_matches_library_circular_conv():
    g = torch.Generator().manual_seed(0)
    x = torch.randn(2, 6, 8, 5, dtype=torch.float64, generator=g)
    w = torch.randn(9, 7, 5, dtype=torch.float64, generator=g)
    b = torch.randn(7, dtype=torch.float64, generator=g)
    mine = convnet.conv3x3(x, w, b)
    # conv2d is a cross-correlation: y[p] = sum_{ky,kx} k[ky, kx] x[p + (ky - 1, kx - 1)], tap t = 3 ky + kx
    k = w.reshape(3, 3, 7, 5).permute(2, 3, 0, 1)
    xp = F.pad(x.permute(0, 3, 1, 2), (1, 1, 1, 1), mode="circular")
    ref = F.conv2d(xp, k, b).permute(0, 2, 3, 1)
    assert torch.allclose(mine, ref, rtol=1e-13, atol=1e-13)


def test_pool_and_upsample_closed_forms():
    x = torch.arange(2 * 4 * 6 * 3, dtype=torch.float64).reshape(2, 4, 6, 3)
    p = convnet.avgpool2(x)
    assert p.shape == (2, 2, 3, 3)
    assert p[1, 1, 2, 0] == (x[1, 2, 4, 0] + x[1, 2, 5, 0] + x[1, 3, 4, 0] + x[1, 3, 5, 0]) / 4
    u = convnet.upsample2(p)
    assert u.shape == x.shape and torch.equal(u[:, 1::2, 1::2], p) and torch.equal(u[:, ::2, 1::2], p)
    # the adjoint of nearest upsampling is the 2x2 block sum: <up(a), b> = <a, 4 avgpool(b)>
    a, bb = torch.randn(2, 2, 3, 3, dtype=torch.float64), torch.randn(2, 4, 6, 3, dtype=torch.float64)
    assert torch.allclose((convnet.upsample2(a) * bb).sum(), (a * 4 * convnet.avgpool2(bb)).sum())


def make_params(channels, seed=0):
    g = torch.Generator().manual_seed(seed)
    L = len(channels)
    shapes = []
    for l in range(L):
        cin = 1 if l == 0 else channels[l - 1]
        shapes += [(channels[l], cin), (channels[l], channels[l])]
    for l in range(L - 1):
        shapes += [(channels[l], channels[l + 1] + channels[l]), (channels[l], channels[l])]
    shapes.append((2, channels[0]))
    params = []
    for co, ci in shapes:
        w = torch.randn(9, co, ci, dtype=torch.float64, generator=g) * np.sqrt(2.0 / (9 * ci))
        b = 0.1 * torch.randn(co, dtype=torch.float64, generator=g)
        params.append((w.requires_grad_(), b.requires_grad_()))
    return params


def test_unet_gradient_finite_differences():
    ch = (4, 6)
    params = make_params(ch)
    g = torch.Generator().manual_seed(1)
    z = torch.randn(1, 8, 8, 1, dtype=torch.float64, generator=g)
    tgt = torch.randn(1, 8, 8, 2, dtype=torch.float64, generator=g)

    def loss():
        return ((convnet.unet_forward(z, params, ch) - tgt) ** 2).mean()

    L = loss()
    L.backward()
    rng = np.random.default_rng(2)
    h = 1e-6
    for li in (0, 3, 4, len(params) - 1):  # first encoder, second level, decoder concat conv, output
        w, b = params[li]
        for _ in range(3):
            idx = tuple(int(rng.integers(0, s)) for s in w.shape)
            with torch.no_grad():
                w[idx] += h
                lp = loss().item()
                w[idx] -= 2 * h
                lm = loss().item()
                w[idx] += h
            fd = (lp - lm) / (2 * h)
            assert fd == pytest.approx(w.grad[idx].item(), rel=1e-5, abs=1e-9)
        with torch.no_grad():
            b[0] += h
            lp = loss().item()
            b[0] -= 2 * h
            lm = loss().item()
            b[0] += h
        assert (lp - lm) / (2 * h) == pytest.approx(b.grad[0].item(), rel=1e-5, abs=1e-9)


def test_conv3x3x3_matches_library_circular_conv3d():
    g = torch.Generator().manual_seed(3)
    x = torch.randn(2, 4, 6, 8, 3, dtype=torch.float64, generator=g)
    w = torch.randn(27, 5, 3, dtype=torch.float64, generator=g)
    b = torch.randn(5, dtype=torch.float64, generator=g)
    mine = convnet.conv3x3x3(x, w, b)
    # conv3d cross-correlation, tap t = 9 kz + 3 ky + kx
    k = w.reshape(3, 3, 3, 5, 3).permute(3, 4, 0, 1, 2)
    xp = F.pad(x.permute(0, 4, 1, 2, 3), (1, 1, 1, 1, 1, 1), mode="circular")
    ref = F.conv3d(xp, k, b).permute(0, 2, 3, 4, 1)
    assert torch.allclose(mine, ref, rtol=1e-13, atol=1e-13)
    p = convnet.avgpool2x2x2(x[:, :, :, :, :1].repeat(1, 1, 1, 1, 2))
    assert p.shape == (2, 2, 3, 4, 2)
    assert torch.allclose(p[0, 1, 2, 3, 0], x[0, 2:4, 4:6, 6:8, 0].mean())
    u = convnet.upsample2x2x2(p)
    assert torch.equal(u[:, 1::2, ::2, 1::2], p)


def test_unet3d_gradient_finite_differences():
    ch = (4, 6)
    g = torch.Generator().manual_seed(4)
    L = len(ch)
    shapes = []
    for l in range(L):
        cin = 1 if l == 0 else ch[l - 1]
        shapes += [(ch[l], cin), (ch[l], ch[l])]
    for l in range(L - 1):
        shapes += [(ch[l], ch[l + 1] + ch[l]), (ch[l], ch[l])]
    shapes.append((3, ch[0]))
    params = []
    for co, ci in shapes:
        w = (torch.randn(27, co, ci, dtype=torch.float64, generator=g) * np.sqrt(2.0 / (27 * ci))).requires_grad_()
        params.append((w, (0.1 * torch.randn(co, dtype=torch.float64, generator=g)).requires_grad_()))
    z = torch.randn(1, 4, 4, 4, 1, dtype=torch.float64, generator=g)
    tgt = torch.randn(1, 4, 4, 4, 3, dtype=torch.float64, generator=g)

    def loss():
        return ((convnet.unet_forward(z, params, ch) - tgt) ** 2).mean()

    loss().backward()
    h = 1e-6
    rng = np.random.default_rng(5)
    for li in (0, 4, len(params) - 1):
        w, _ = params[li]
        idx = tuple(int(rng.integers(0, s)) for s in w.shape)
        with torch.no_grad():
            w[idx] += h
            lp = loss().item()
            w[idx] -= 2 * h
            lm = loss().item()
            w[idx] += h
        assert (lp - lm) / (2 * h) == pytest.approx(w.grad[idx].item(), rel=1e-5, abs=1e-9)


def test_mse_multiscale_closed_forms():
    # e constant: the centred cell means vanish, L = sum e^2
    B, H, W, C = 2, 8, 8, 2
    z = np.zeros((B, H, W, C))
    L, g = convnet.mse_multiscale(z + 0.5, z, 3.0, 3)
    assert abs(L - 0.25 * z.size) < 1e-12 and torch.allclose(g, torch.full_like(g, 1.0))
    # finest checkerboard: every cell mean is zero, L = sum e^2 = N
    y, x = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    chk = ((-1.0) ** (x + y))[None, :, :, None] * np.ones((B, 1, 1, C))
    L, _ = convnet.mse_multiscale(chk, z, 3.0, 3)
    assert abs(L - z.size) < 1e-12
    # +-1 constant on 2x2 cells in a checkerboard of cells: level-1 means +-1, coarser ones zero,
    # so L = N + lam 4 (N / 4) = N (1 + lam)
    chk2 = ((-1.0) ** (x // 2 + y // 2))[None, :, :, None] * np.ones((B, 1, 1, C))
    lam = 3.0
    L, g = convnet.mse_multiscale(chk2, z, lam, 3)
    assert abs(L - z.size * (1.0 + lam)) < 1e-9
    assert torch.allclose(g, torch.as_tensor(2.0 * (1.0 + lam) * chk2))
    # 3-D: +-1 on 2x2x2 cells in a 3-D checkerboard: L = N (1 + lam)
    Z = 4
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(4), np.arange(4), indexing="ij")
    c3 = ((-1.0) ** (xx // 2 + yy // 2 + zz // 2))[None, ..., None] * np.ones((1, 1, 1, 1, 3))
    L, _ = convnet.mse_multiscale(c3, np.zeros_like(c3), lam, 2, dim=3)
    assert abs(L - c3.size * (1.0 + lam)) < 1e-9
    # lam = 0 is the plain squared error
    rng = np.random.default_rng(0)
    p, t = rng.standard_normal((B, H, W, C)), rng.standard_normal((B, H, W, C))
    L, g = convnet.mse_multiscale(p, t, 0.0, 3)
    assert abs(L - ((p - t) ** 2).sum()) < 1e-9 and torch.allclose(g, torch.as_tensor(2.0 * (p - t)))


@pytest.mark.parametrize("dim", [2, 3])
def test_mse_multiscale_gradient_finite_differences(dim):
    rng = np.random.default_rng(1)
    shape = (2, 8, 8, 2) if dim == 2 else (1, 4, 4, 4, 3)
    p, t = rng.standard_normal(shape), rng.standard_normal(shape)
    lam, levels = 0.7, (3 if dim == 2 else 2)
    _, g = convnet.mse_multiscale(p, t, lam, levels, dim=dim)
    h = 1e-6
    for k in rng.choice(p.size, 12, replace=False):
        pp, pm = p.copy(), p.copy()
        pp.flat[k] += h
        pm.flat[k] -= h
        fd = (convnet.mse_multiscale(pp, t, lam, levels, dim=dim)[0] - convnet.mse_multiscale(pm, t, lam, levels, dim=dim)[0]) / (2 * h)
        assert abs(fd - g.flatten()[k].item()) < 1e-6 * max(1.0, abs(fd))
