"""Learned warm start for the tied PCG of a periodic Quad4 cell (no counterpart in the reference).

A periodic encoder-decoder on the element grid maps the modulus field log G (one value per
element) to the displacement fluctuation of the independent nodes for a unit macroscopic shear
(the solution is linear in the load increment).  The prediction, plus the affine part of the
load, is the initial guess of the FP64 Jacobi-PCG on K' (MatrixPartitionedTyings.h:655-674).
The CG then runs to the same true-residual tolerance as from a zero guess, so the warm start
changes the iteration count and never the accuracy.

Every operation of training and inference runs in the library's HIP kernels (gf_conv.hip):
BF16 MFMA 3x3 circular convolutions in NHWC, FP32 accumulation, FP32 master weights and Adam,
deterministic reductions.  torch provides device memory and the stream.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _call, _dp, _stream, _vp, _i64, _f64, _i32, randn, to_device, device

_f32 = ctypes.c_float
CF_RELU, CF_ACCUM = 1, 2
IN_CH = 16   # network input channels in storage (channel 0 = normalised log G)
OUT_CH = 16  # output-gradient channels in storage (channels 0, 1 = the two displacement components)


def _torch():
    import torch

    return torch


def coutp_of(c):
    for p in (32, 64, 128, 256):
        if c <= p:
            return p
    raise ValueError("at most 256 channels")


class ConvLayer:
    """3x3 circular convolution: FP32 master weights [9][cout][cin] + bias, Adam state, packed BF16
    forward weights [9][coutp][cin_s] and data-gradient weights [9][cinp][cout_s]."""

    def __init__(self, cin, cout, cin_s, cout_s, seed):
        torch = _torch()
        self.cin, self.cout, self.cin_s, self.cout_s = cin, cout, cin_s, cout_s
        self.coutp = coutp_of(cout_s)
        self.cinp = coutp_of(cin_s)
        std = np.sqrt(2.0 / (9 * cin))  # He initialisation
        w = randn(seed, 9 * cout * cin, 0.0, std).astype(np.float32)
        self.w = to_device(w, torch.float32)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device())
        self.dw = torch.zeros_like(self.w)
        self.db = torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.wf = torch.zeros(9 * self.coutp * cin_s, dtype=torch.bfloat16, device=device())
        self.wb = torch.zeros(9 * self.cinp * cout_s, dtype=torch.bfloat16, device=device())
        self.pack()

    def pack(self):
        _call("gf_conv_pack", _i32(self.cout), _i32(self.cin), _dp(self.w), _i32(self.coutp), _i32(self.cin_s),
              _dp(self.wf), _i32(self.cinp), _i32(self.cout_s), _dp(self.wb), _stream())

    def adam(self, lr, step, b1=0.9, b2=0.999, eps=1e-8):
        for w, g, m, v in ((self.w, self.dw, self.mw, self.vw), (self.b, self.db, self.mb, self.vb)):
            _call("gf_adam", _i64(w.numel()), _dp(w), _dp(g), _dp(m), _dp(v), _f32(lr), _f32(b1), _f32(b2),
                  _f32(eps), _i64(step), _stream())
        self.pack()


def conv(x1, c1, P, H, W, layer, y, flags=0, up1=0, x2=None, c2=0, ysplit=None, y2=None, y32=None, mask=None,
         bias=True, backward=False):
    """y = conv(x) with `layer` (forward weights), or its data gradient (backward=True: tap-flipped,
    transposed weights, output channels = the layer's input channels)."""
    if backward:
        wpk, coutp, cout = layer.wb, layer.cinp, layer.cin_s
    else:
        wpk, coutp, cout = layer.wf, layer.coutp, (layer.cout if y is None else layer.cout_s)
    b = _dp(layer.b) if (bias and not backward) else None
    _call("gf_conv3x3", _dp(x1), _i32(c1), _i32(up1), _dp(x2), _i32(c2), _i64(P), _i32(H), _i32(W), _dp(wpk),
          _i32(coutp), b, _i32(cout), _i32(flags), _dp(y), _i32(ysplit if ysplit is not None else cout), _dp(y2),
          _dp(y32), _dp(mask), _stream())


class WarmStartNet:
    """Periodic U-Net: per level two conv+ReLU (encoder), 2x2 average pooling between levels,
    nearest upsampling + skip concatenation + two conv+ReLU (decoder), final conv to 2 channels.

    channels: per-level widths (multiples of 16, at most 128).  Grid H x W must be divisible by
    2^(levels-1)."""

    def __init__(self, channels=(16, 32, 64, 128), seed=7):
        self.ch = list(channels)
        L = len(self.ch)
        assert all(c % 16 == 0 and c <= 128 for c in self.ch)
        s = seed * 1000
        self.enc, self.dec = [], [None] * max(L - 1, 0)
        for l in range(L):
            cin, cin_s = (1, IN_CH) if l == 0 else (self.ch[l - 1], self.ch[l - 1])
            self.enc.append((ConvLayer(cin, self.ch[l], cin_s, self.ch[l], s + 2 * l),
                             ConvLayer(self.ch[l], self.ch[l], self.ch[l], self.ch[l], s + 2 * l + 1)))
        for l in range(L - 2, -1, -1):
            cin = self.ch[l + 1] + self.ch[l]
            self.dec[l] = (ConvLayer(cin, self.ch[l], cin, self.ch[l], s + 100 + 2 * l),
                           ConvLayer(self.ch[l], self.ch[l], self.ch[l], self.ch[l], s + 101 + 2 * l))
        self.out = ConvLayer(self.ch[0], 2, self.ch[0], OUT_CH, s + 500)
        self.step_count = 0
        self._buf = {}
        self.meta = {}

    def layers(self):
        out = [c for pair in self.enc for c in pair]
        out += [c for pair in self.dec for c in pair]
        return out + [self.out]

    def nparams(self):
        return sum(c.w.numel() + c.b.numel() for c in self.layers())

    # ---- buffers ----
    def buffers(self, B, H, W):
        key = (B, H, W)
        if key in self._buf:
            return self._buf[key]
        torch = _torch()
        L = len(self.ch)
        assert H % (1 << (L - 1)) == 0 and W % (1 << (L - 1)) == 0, "grid not divisible by the pooling depth"

        def t(P, C):
            return torch.zeros(P * C, dtype=torch.bfloat16, device=device())

        lv = [(B * (H >> l) * (W >> l), H >> l, W >> l) for l in range(L)]
        bf = {"lv": lv, "X0": t(lv[0][0], IN_CH), "pred": torch.zeros(lv[0][0] * 2, dtype=torch.float32, device=device()),
              "gY": t(lv[0][0], OUT_CH), "I": [None] * L, "A": [], "E": [], "Bd": [None] * L, "D": [None] * L,
              "gI": [None] * L, "gA": [], "gE": [], "gB": [None] * L, "gD": [None] * L, "gUp": [None] * L}
        for l in range(L):
            P = lv[l][0]
            if l > 0:
                bf["I"][l] = t(P, self.ch[l - 1])
                bf["gI"][l] = t(P, self.ch[l - 1])
            bf["A"].append(t(P, self.ch[l]))
            bf["E"].append(t(P, self.ch[l]))
            bf["gA"].append(t(P, self.ch[l]))
            bf["gE"].append(t(P, self.ch[l]))
            if l < L - 1:
                bf["Bd"][l] = t(P, self.ch[l])
                bf["D"][l] = t(P, self.ch[l])
                bf["gB"][l] = t(P, self.ch[l])
                bf["gD"][l] = t(P, self.ch[l])
                bf["gUp"][l] = t(P, self.ch[l + 1])
        # weight-gradient workspace: the largest layer
        need = 0
        for c in self.layers():
            lvl = self._level_of(c)
            ns, nf = ctypes.c_int(0), ctypes.c_int64(0)
            _call("gf_conv3x3_wgrad_work", _i64(lv[lvl][0]), _i32(c.cout_s if c is not self.out else OUT_CH),
                  _i32(c.cin_s), ctypes.byref(ns), ctypes.byref(nf))
            need = max(need, nf.value)
        bf["work"] = torch.zeros(need, dtype=torch.float32, device=device())
        bf["loss_part"] = torch.zeros((lv[0][0] * OUT_CH + 255) // 256, dtype=torch.float32, device=device())
        bf["loss"] = torch.zeros(1, dtype=torch.float32, device=device())
        self._buf[key] = bf
        return bf

    def _level_of(self, layer):
        for l, pair in enumerate(self.enc):
            if layer in pair:
                return l
        for l, pair in enumerate(self.dec):
            if layer in pair:
                return l
        return 0

    def _wgrad(self, layer, bf, l, x1, c1, g, up1=0, x2=None, c2=0):
        P, H, W = bf["lv"][l]
        work = bf["work"]
        cg = layer.cout_s
        _call("gf_conv3x3_wgrad", _dp(x1), _i32(c1), _i32(up1), _dp(x2), _i32(c2), _i64(P), _i32(H), _i32(W), _dp(g),
              _i32(cg), _i32(layer.cout), _i32(layer.cin), _dp(work), _i64(work.numel()), _dp(layer.dw), _dp(layer.db),
              _stream())

    # ---- forward / backward ----
    def forward(self, B, H, W):
        """Network on the input already in buffers(B, H, W)["X0"]; returns the FP32 prediction [B*H*W*2]."""
        bf = self.buffers(B, H, W)
        L = len(self.ch)
        x, cx = bf["X0"], IN_CH
        for l in range(L):
            P, h, w = bf["lv"][l]
            if l > 0:
                _call("gf_avgpool2", _dp(bf["E"][l - 1]), _i64(B), _i32(2 * h), _i32(2 * w), _i32(self.ch[l - 1]),
                      _dp(bf["I"][l]), _stream())
                x, cx = bf["I"][l], self.ch[l - 1]
            c1, c2 = self.enc[l]
            conv(x, cx, P, h, w, c1, bf["A"][l], flags=CF_RELU)
            conv(bf["A"][l], self.ch[l], P, h, w, c2, bf["E"][l], flags=CF_RELU)
        d, cd = bf["E"][L - 1], self.ch[L - 1]
        for l in range(L - 2, -1, -1):
            P, h, w = bf["lv"][l]
            c1, c2 = self.dec[l]
            conv(d, cd, P, h, w, c1, bf["Bd"][l], flags=CF_RELU, up1=1, x2=bf["E"][l], c2=self.ch[l])
            conv(bf["Bd"][l], self.ch[l], P, h, w, c2, bf["D"][l], flags=CF_RELU)
            d, cd = bf["D"][l], self.ch[l]
        P, h, w = bf["lv"][0]
        conv(d, cd, P, h, w, self.out, None, y32=bf["pred"])
        return bf["pred"]

    def backward(self, B, H, W):
        """Gradients of all layers from bf["gY"] (dL/dpred in channels 0, 1)."""
        bf = self.buffers(B, H, W)
        L = len(self.ch)
        P0, h0, w0 = bf["lv"][0]
        last = bf["D"][0] if L > 1 else bf["E"][0]
        self._wgrad(self.out, bf, 0, last, self.ch[0], bf["gY"])
        g_last = bf["gD"][0] if L > 1 else bf["gE"][0]
        conv(bf["gY"], OUT_CH, P0, h0, w0, self.out, g_last, mask=last, backward=True)
        for l in range(L - 1):
            P, h, w = bf["lv"][l]
            c1, c2 = self.dec[l]
            self._wgrad(c2, bf, l, bf["Bd"][l], self.ch[l], bf["gD"][l])
            conv(bf["gD"][l], self.ch[l], P, h, w, c2, bf["gB"][l], mask=bf["Bd"][l], backward=True)
            dn = bf["D"][l + 1] if l + 1 < L - 1 else bf["E"][L - 1]
            self._wgrad(c1, bf, l, dn, self.ch[l + 1], bf["gB"][l], up1=1, x2=bf["E"][l], c2=self.ch[l])
            conv(bf["gB"][l], self.ch[l], P, h, w, c1, bf["gUp"][l], ysplit=self.ch[l + 1], y2=bf["gE"][l],
                 backward=True)
            gdn = bf["gD"][l + 1] if l + 1 < L - 1 else bf["gE"][L - 1]
            _call("gf_sumpool2", _dp(bf["gUp"][l]), _i64(B), _i32(h), _i32(w), _i32(self.ch[l + 1]), _dp(dn),
                  _dp(gdn), _stream())
        for l in range(L - 1, -1, -1):
            P, h, w = bf["lv"][l]
            if l < L - 1:
                # skip gradient from the decoder + the pooled path of level l + 1, masked by E_l > 0
                _call("gf_unpool2", _dp(bf["gI"][l + 1]), _i64(B), _i32(h), _i32(w), _i32(self.ch[l]), _f32(0.25),
                      _i32(1), _dp(bf["E"][l]), _dp(bf["gE"][l]), _stream())
            c1, c2 = self.enc[l]
            self._wgrad(c2, bf, l, bf["A"][l], self.ch[l], bf["gE"][l])
            conv(bf["gE"][l], self.ch[l], P, h, w, c2, bf["gA"][l], mask=bf["A"][l], backward=True)
            x, cx = (bf["X0"], IN_CH) if l == 0 else (bf["I"][l], self.ch[l - 1])
            self._wgrad(c1, bf, l, x, cx, bf["gA"][l])
            if l > 0:
                conv(bf["gA"][l], self.ch[l], P, h, w, c1, bf["gI"][l], backward=True)

    def loss_grad(self, B, H, W, target, scale=1.0):
        """MSE of the prediction against target [B*H*W*2] FP32; writes gY; returns the device sum of
        squared errors (a 1-element tensor)."""
        bf = self.buffers(B, H, W)
        P0 = bf["lv"][0][0]
        _call("gf_mse", _i64(P0), _i32(2), _dp(bf["pred"]), _dp(target), _f32(scale), _i32(OUT_CH), _dp(bf["gY"]),
              _dp(bf["loss_part"]), _dp(bf["loss"]), _stream())
        return bf["loss"]

    def adam(self, lr):
        self.step_count += 1
        for c in self.layers():
            c.adam(lr, self.step_count)

    def set_input(self, B, H, W, G_list, shift, scale):
        """Normalised log G of B element fields (device FP64 [H*W] each) into the input buffer."""
        bf = self.buffers(B, H, W)
        HW = H * W
        X0 = bf["X0"]
        for b, G in enumerate(G_list):
            xb = X0[b * HW * IN_CH:(b + 1) * HW * IN_CH]
            _call("gf_input_logG", _i64(HW), _dp(G), _f64(shift), _f64(scale), _i32(IN_CH), _dp(xb), _stream())

    # ---- persistence ----
    def state(self):
        d = {"channels": np.array(self.ch, dtype=np.int64)}
        for i, c in enumerate(self.layers()):
            d[f"w{i}"] = c.w.cpu().numpy()
            d[f"b{i}"] = c.b.cpu().numpy()
        for k, v in self.meta.items():
            d["meta_" + k] = np.asarray(v)
        return d

    def save(self, path):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path):
        torch = _torch()
        with np.load(path, allow_pickle=False) as f:
            net = cls(tuple(int(c) for c in f["channels"]))
            for i, c in enumerate(net.layers()):
                c.w.copy_(torch.from_numpy(f[f"w{i}"]))
                c.b.copy_(torch.from_numpy(f[f"b{i}"]))
                c.pack()
            net.meta = {k[5:]: f[k][()] if f[k].ndim == 0 else f[k] for k in f.files if k.startswith("meta_")}
        return net


class PeriodicShearWarmStart:
    """Couples a WarmStartNet to the tied PCG of a periodic Quad4 Regular(n, n) cell under a simple
    shear increment of the control DOF ctrl(0, 1) (docs/examples/statics/periodic_elastoplastic.py:141).

    The independent nodes (i, j), 0 <= i, j < n, are the lower-left nodes of the elements (i, j), so
    the nodal grid and the element grid coincide.  Unknown x_u of node (i, j), component c:
    gamma * (u_fluct(i, j, c) + (c == 0) * y_j); the origin node is prescribed."""

    def __init__(self, pcg, tyings, n, net, sigma=0.4, target_scale=None):
        torch = _torch()
        self.P, self.ty, self.n, self.net, self.sigma = pcg, tyings, n, net, sigma
        dofs = np.asarray(tyings.dofs)
        idx = np.full((n, n, 2), -1, dtype=np.int32)
        aff = np.zeros((n, n, 2))
        for j in range(n):
            nodes = j * (n + 1) + np.arange(n)
            d = dofs[nodes]
            idx[j] = np.where(d < tyings.nnu, d, -1)
            aff[j, :, 0] = float(j)  # affine u_x = gamma * y for ctrl(0, 1) = gamma
        self.idx = to_device(idx.ravel(), torch.int32)
        self.affine = to_device(aff.ravel())
        self.xp_unit = np.zeros(tyings.nnp)
        self.xp_unit[tyings.control[0, 1] - tyings.nnu] = 1.0
        self.target_scale = float(target_scale if target_scale is not None else net.meta.get("target_scale", 1.0))

    def rhs(self, gamma):
        return -self.P.apply(None, gamma * self.xp_unit)

    def predict(self, G, gamma, out=None):
        """Initial guess x_u (device FP64 [nnu]) for the modulus field G (device FP64 [n*n])."""
        torch = _torch()
        n = self.n
        self.net.set_input(1, n, n, [G], 0.0, 1.0 / self.sigma)
        pred = self.net.forward(1, n, n)
        x = out if out is not None else torch.zeros(self.ty.nnu, dtype=torch.float64, device=device())
        _call("gf_grid_to_dofs", _i64(n * n * 2), _i32(2), _dp(pred), _dp(self.idx), _dp(self.affine),
              _f64(self.target_scale), _f64(gamma), _dp(x), _stream())
        return x

    def target(self, x_u, gamma, out):
        """Centred, scaled fluctuation grid [n*n*2] FP32 of a solution x_u."""
        n = self.n
        _call("gf_dofs_to_grid", _i64(n * n * 2), _i32(2), _dp(x_u), _dp(self.idx), _dp(self.affine),
              _f64(self.target_scale), _f64(gamma), _i32(1), _dp(out), _stream())
        return out

    def make_dataset(self, K, seeds, tol=1e-6, progress=None):
        """PCG solutions for the modulus fields exp(sigma z_s); returns (G [ns, n*n] FP64,
        targets [ns, n*n*2] FP32, total iterations, seconds)."""
        torch = _torch()
        n = self.n
        ns = len(seeds)
        Gs = torch.zeros((ns, n * n), dtype=torch.float64, device=device())
        T = torch.zeros((ns, n * n * 2), dtype=torch.float32, device=device())
        Kd = to_device(K)
        its = 0
        t0 = time.perf_counter()
        for k, s in enumerate(seeds):
            Gs[k] = to_device(np.exp(self.sigma * randn(s, n * n)))
            self.P.set_moduli(Kd, Gs[k])
            b = self.rhs(1.0)
            x, it, _ = self.P.solve(b, tol=tol)
            its += it
            self.target(x, 1.0, T[k])
            if progress and (k % progress == 0 or k == ns - 1):
                print(f"dataset {k + 1}/{ns}: {it} iterations, {time.perf_counter() - t0:.1f} s", flush=True)
        torch.cuda.synchronize()
        return Gs, T, its, time.perf_counter() - t0

    def holdout_error(self, Gs, T):
        """Mean relative L2 error of the prediction (up to a translation) against centred targets."""
        n = self.n
        errs = []
        for k in range(Gs.shape[0]):
            self.net.set_input(1, n, n, [Gs[k]], 0.0, 1.0 / self.sigma)
            p = self.net.forward(1, n, n).reshape(n * n, 2)
            p = p - p.mean(dim=0, keepdim=True)
            t = T[k].reshape(n * n, 2)
            errs.append(float(((p - t).norm() / t.norm()).item()))
        return float(np.mean(errs))

    def train(self, Gs, T, epochs, batch=4, lr=2e-3, lr_min=1e-5, log_every=1, time_limit=None):
        """Adam with cosine learning-rate decay on (Gs, T); returns per-epoch mean losses."""
        torch = _torch()
        n = self.n
        ns = Gs.shape[0]
        nb = ns // batch
        total = epochs * nb
        rng = np.random.default_rng(11)
        net = self.net
        bf = net.buffers(batch, n, n)
        tgt = torch.zeros(batch * n * n * 2, dtype=torch.float32, device=device())
        hist = []
        t0 = time.perf_counter()
        step = 0
        for ep in range(epochs):
            perm = rng.permutation(ns)
            acc = torch.zeros(1, dtype=torch.float64, device=device())
            for bi in range(nb):
                ids = perm[bi * batch:(bi + 1) * batch]
                net.set_input(batch, n, n, [Gs[i] for i in ids], 0.0, 1.0 / self.sigma)
                tgt.copy_(T[torch.from_numpy(ids).to(device())].reshape(-1))
                net.forward(batch, n, n)
                loss = net.loss_grad(batch, n, n, tgt)
                acc += loss.double()
                net.backward(batch, n, n)
                lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * step / max(total - 1, 1)))
                net.adam(float(lr_t))
                step += 1
            mean = float(acc.item()) / (nb * batch * n * n * 2)
            hist.append(mean)
            if log_every and (ep % log_every == 0 or ep == epochs - 1):
                print(f"epoch {ep}: mse {mean:.4e}  {time.perf_counter() - t0:.1f} s", flush=True)
            if time_limit and time.perf_counter() - t0 > time_limit:
                break
        del bf
        return hist


def periodic_problem(n, sigma=0.4, seed=4, K=10.0):
    """Periodic Quad4 Regular(n, n) cell with Tyings::Control + Tyings::Periodic as in
    docs/examples/statics/periodic_elastoplastic.py:28-47 (iip = control DOFs + origin DOFs),
    plane strain, per-element G = exp(sigma z), K constant; returns a namespace with the tied PCG."""
    from types import SimpleNamespace

    import goosefem_amd as gf

    mesh = gf.Mesh.Quad4.Regular(n, n)
    ctrl = gf.Tyings.Control(mesh.coor, mesh.dofs)
    iip = np.concatenate((ctrl.controlDofs.ravel(), ctrl.dofs[mesh.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, mesh.nodesPeriodic, iip)
    vec = gf.Vector(mesh.conn, ty.dofs)
    quad = gf.Element.Quad4.QuadraturePlanar(gf.Vector(mesh.conn, mesh.dofs).AsElement(mesh.coor))
    G = np.exp(sigma * randn(seed, mesh.nelem))
    Kv = np.full(mesh.nelem, K)
    P = gf.TyingsPCG(vec, quad, Kv, G, ty)
    return SimpleNamespace(mesh=mesh, ty=ty, vec=vec, quad=quad, G=G, K=Kv, P=P, n=n, sigma=sigma)


def train_main(argv=None):
    """Generate PCG solutions for random modulus fields, train one or more networks on the device,
    report their error on held-out fields, save the weights (npz) with the measured data-generation
    and training times."""
    import argparse
    import os

    torch = _torch()
    ap = argparse.ArgumentParser(description="train the learned PCG warm start (periodic Quad4 n x n, simple shear)")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--samples", type=int, default=400)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", default="2e-3", help="learning rate (';'-separated, one per configuration)")
    ap.add_argument("--channels", default="16,32,64,128", help="per-level widths; ';' separates configurations")
    ap.add_argument("--data-tol", type=float, default=1e-7)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--time-limit", type=float, default=None, help="seconds of training per configuration at most")
    ap.add_argument("--out", required=True, help="weights file (.npz); with several configurations a suffix is added")
    a = ap.parse_args(argv)
    pb = periodic_problem(a.n)
    configs = [tuple(int(c) for c in cfg.split(",")) for cfg in a.channels.split(";")]
    lrs = [float(x) for x in a.lr.split(";")]
    lrs = lrs + [lrs[-1]] * (len(configs) - len(lrs))
    probe = PeriodicShearWarmStart(pb.P, pb.ty, a.n, WarmStartNet(configs[0]), sigma=pb.sigma, target_scale=1.0)
    seeds = list(range(a.seed0, a.seed0 + a.samples + a.holdout))
    Gs, T, its, t_data = probe.make_dataset(pb.K, seeds, tol=a.data_tol, progress=50)
    scale = float(T[: a.samples].std().item())
    T.div_(scale)
    print(f"data: {len(seeds)} solves, {its} CG iterations, {t_data:.1f} s; target scale {scale:.4e}", flush=True)
    Gtr, Ttr, Gho, Tho = Gs[: a.samples], T[: a.samples], Gs[a.samples:], T[a.samples:]
    for cfg, lr in zip(configs, lrs):
        net = WarmStartNet(cfg)
        W = PeriodicShearWarmStart(pb.P, pb.ty, a.n, net, sigma=pb.sigma, target_scale=scale)
        t0 = time.perf_counter()
        hist = W.train(Gtr, Ttr, a.epochs, batch=a.batch, lr=lr, time_limit=a.time_limit)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        err = W.holdout_error(Gho, Tho) if a.holdout else float("nan")
        net.meta = {"target_scale": scale, "sigma": pb.sigma, "n": a.n, "samples": a.samples, "epochs": len(hist),
                    "batch": a.batch, "lr": lr, "data_seconds": t_data, "train_seconds": t_train, "final_mse": hist[-1],
                    "holdout_rel_error": err, "data_tol": a.data_tol, "seed0": a.seed0, "params": net.nparams()}
        out = a.out
        if len(configs) > 1:
            stem, ext = os.path.splitext(a.out)
            out = f"{stem}_{'-'.join(map(str, cfg))}{ext}"
        net.save(out)
        print(f"channels {cfg} lr {lr}: {net.nparams()} parameters, {len(hist)} epochs in {t_train:.1f} s; final mse "
              f"{hist[-1]:.4e}; held-out relative error {err:.4f}; saved {out}", flush=True)


if __name__ == "__main__":
    train_main()
