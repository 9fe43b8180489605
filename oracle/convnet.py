"""FP64 CPU restatement of the warm-start network (test infrastructure only: tests/ compare
goosefem_amd/csrc/gf_conv.hip against it; nothing in goosefem_amd/ imports it).

Same layout and architecture as the device implementation: NHWC activations, weights
[9][cout][cin] with tap t = 3 (dy + 1) + (dx + 1), 3x3 circular convolution, 2x2 average pooling,
nearest upsampling with the upsampled tensor first in the skip concatenation, ReLU after every
convolution but the last.  Gradients come from torch autograd in float64 and are pinned by a
central finite-difference check in tests/test_oracle_convnet.py.  No counterpart in the reference.
"""

import torch


def conv3x3x3(x, w, b=None):
    """x [N, Z, H, W, Cin], w [27, Cout, Cin] -> [N, Z, H, W, Cout];
    y[p] = sum_t w[t] x[p + (dz, dy, dx)], t = 9 (dz + 1) + 3 (dy + 1) + dx + 1."""
    out = None
    for t in range(27):
        dz, dy, dx = t // 9 - 1, (t % 9) // 3 - 1, t % 3 - 1
        xs = torch.roll(x, shifts=(-dz, -dy, -dx), dims=(1, 2, 3))
        term = torch.einsum("nzhwc,oc->nzhwo", xs, w[t])
        out = term if out is None else out + term
    if b is not None:
        out = out + b
    return out


def avgpool2x2x2(x):
    N, Z, H, W, C = x.shape
    return x.reshape(N, Z // 2, 2, H // 2, 2, W // 2, 2, C).mean(dim=(2, 4, 6))


def upsample2x2x2(x):
    return x.repeat_interleave(2, dim=1).repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)


def conv3x3(x, w, b=None):
    """x [N, H, W, Cin], w [9, Cout, Cin] -> [N, H, W, Cout]; y[p] = sum_t w[t] x[p + (dy, dx)]."""
    out = None
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        xs = torch.roll(x, shifts=(-dy, -dx), dims=(1, 2))
        term = torch.einsum("nhwc,oc->nhwo", xs, w[t])
        out = term if out is None else out + term
    if b is not None:
        out = out + b
    return out


def avgpool2(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(dim=(2, 4))


def upsample2(x):
    return x.repeat_interleave(2, dim=1).repeat_interleave(2, dim=2)


def bf16(x):
    """Round to bfloat16 (round to nearest even) and back."""
    return x.to(torch.bfloat16).to(x.dtype)


def unet_forward(z, params, channels, round_bf16=False):
    """z [N, H, W, 1] or [N, Z, H, W, 1] (normalised log G); params: list of (w, b) in the device
    network's layer order (encoder pairs per level, decoder pairs per level 0..L-2, output layer).
    With round_bf16 the activations are rounded to bfloat16 where the device stores them."""
    L = len(channels)
    if z.dim() == 5:
        conv, pool, up = conv3x3x3, avgpool2x2x2, upsample2x2x2
    else:
        conv, pool, up = conv3x3, avgpool2, upsample2
    r = bf16 if round_bf16 else (lambda t: t)
    enc = params[: 2 * L]
    dec = params[2 * L: 2 * L + 2 * (L - 1)]
    out = params[-1]
    x = r(z)
    E = []
    for l in range(L):
        if l > 0:
            x = r(pool(E[-1]))
        (w1, b1), (w2, b2) = enc[2 * l], enc[2 * l + 1]
        a = r(torch.relu(conv(x, w1, b1)))
        e = r(torch.relu(conv(a, w2, b2)))
        E.append(e)
    d = E[-1]
    for l in range(L - 2, -1, -1):
        (w1, b1), (w2, b2) = dec[2 * l], dec[2 * l + 1]
        cat = torch.cat([up(d), E[l]], dim=-1)
        bb = r(torch.relu(conv(cat, w1, b1)))
        d = r(torch.relu(conv(bb, w2, b2)))
    return conv(d, out[0], out[1])


def mse_multiscale(pred, target, lam, levels, dim=2):
    """Multi-scale mean squared error of the warm start's training (gf_mse_multiscale), float64.
    pred, target [B, H, W, C] (dim 2) or [B, Z, H, W, C] (dim 3).  With e = pred - target, mu each
    grid's per-channel mean of e and m_s the mean of e over cells 2^s wide (s = 1..levels):
        L = sum e^2 + lam sum_s (2^dim)^s sum_cells (m_s - mu)^2.
    Returns (L, dL/de) with the gradient from autograd; pinned by finite differences and closed
    forms in tests/test_oracle_convnet.py."""
    e = (torch.as_tensor(pred, dtype=torch.float64) - torch.as_tensor(target, dtype=torch.float64)).clone()
    e.requires_grad_(True)
    sp = tuple(range(1, dim + 1))
    mu = e.mean(dim=sp, keepdim=True)
    loss = (e * e).sum()
    for s in range(1, levels + 1):
        f = 1 << s
        if dim == 2:
            B, H, W, C = e.shape
            m = e.reshape(B, H // f, f, W // f, f, C).mean(dim=(2, 4))
        else:
            B, Z, H, W, C = e.shape
            m = e.reshape(B, Z // f, f, H // f, f, W // f, f, C).mean(dim=(2, 4, 6))
        loss = loss + lam * float((1 << dim) ** s) * ((m - mu) ** 2).sum()
    (g,) = torch.autograd.grad(loss, e)
    return float(loss.detach()), g.detach()
