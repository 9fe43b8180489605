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
    """3x3 (9 taps) or 3x3x3 (27 taps) circular convolution: FP32 master weights [ntap][cout][cin] +
    bias, Adam state, packed BF16 forward weights [ntap][coutp][cin_s] and data-gradient weights
    [ntap][cinp][cout_s]."""

    def __init__(self, cin, cout, cin_s, cout_s, seed, ntap=9):
        torch = _torch()
        self.cin, self.cout, self.cin_s, self.cout_s, self.ntap = cin, cout, cin_s, cout_s, ntap
        self.coutp = coutp_of(cout_s)
        self.cinp = coutp_of(cin_s)
        std = np.sqrt(2.0 / (ntap * cin))  # He initialisation
        w = randn(seed, ntap * cout * cin, 0.0, std).astype(np.float32)
        self.w = to_device(w, torch.float32)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device())
        self.dw = torch.zeros_like(self.w)
        self.db = torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.wf = torch.zeros(ntap * self.coutp * cin_s, dtype=torch.bfloat16, device=device())
        self.wb = torch.zeros(ntap * self.cinp * cout_s, dtype=torch.bfloat16, device=device())
        self.pack()

    def pack(self):
        _call("gf_conv_pack3" if self.ntap == 27 else "gf_conv_pack", _i32(self.cout), _i32(self.cin), _dp(self.w),
              _i32(self.coutp), _i32(self.cin_s), _dp(self.wf), _i32(self.cinp), _i32(self.cout_s), _dp(self.wb),
              _stream())

    def adam(self, lr, step, b1=0.9, b2=0.999, eps=1e-8):
        for w, g, m, v in ((self.w, self.dw, self.mw, self.vw), (self.b, self.db, self.mb, self.vb)):
            _call("gf_adam", _i64(w.numel()), _dp(w), _dp(g), _dp(m), _dp(v), _f32(lr), _f32(b1), _f32(b2),
                  _f32(eps), _i64(step), _stream())
        self.pack()


def conv(x1, c1, P, H, W, layer, y, flags=0, up1=0, x2=None, c2=0, ysplit=None, y2=None, y32=None, mask=None,
         bias=True, backward=False, Z=1):
    """y = conv(x) with `layer` (forward weights), or its data gradient (backward=True: tap-flipped,
    transposed weights, output channels = the layer's input channels).  3-D layers take the depth Z."""
    if backward:
        wpk, coutp, cout = layer.wb, layer.cinp, layer.cin_s
    else:
        wpk, coutp, cout = layer.wf, layer.coutp, (layer.cout if y is None else layer.cout_s)
    b = _dp(layer.b) if (bias and not backward) else None
    ys = _i32(ysplit if ysplit is not None else cout)
    if layer.ntap == 27:
        _call("gf_conv3x3x3", _dp(x1), _i32(c1), _i32(up1), _dp(x2), _i32(c2), _i64(P), _i32(Z), _i32(H), _i32(W),
              _dp(wpk), _i32(coutp), b, _i32(cout), _i32(flags), _dp(y), ys, _dp(y2), _dp(y32), _dp(mask), _stream())
    else:
        _call("gf_conv3x3", _dp(x1), _i32(c1), _i32(up1), _dp(x2), _i32(c2), _i64(P), _i32(H), _i32(W), _dp(wpk),
              _i32(coutp), b, _i32(cout), _i32(flags), _dp(y), ys, _dp(y2), _dp(y32), _dp(mask), _stream())


class WarmStartNet:
    """Periodic U-Net on a 2-D (dim = 2) or 3-D (dim = 3) element grid: per level two conv+ReLU
    (encoder), 2x2(x2) average pooling between levels, nearest upsampling + skip concatenation + two
    conv+ReLU (decoder), final conv to `nout` channels.

    channels: per-level widths (multiples of 16, at most 128).  The grid must be divisible by
    2^(levels-1) in every dimension.  Grids are (H, W) with Z = 1, or (Z, H, W) given as H, W, Z=."""

    def __init__(self, channels=(16, 32, 64, 128), seed=7, dim=2, nout=None):
        self.ch = list(channels)
        self.dim = dim
        self.nout = nout if nout is not None else dim
        ntap = 27 if dim == 3 else 9
        L = len(self.ch)
        assert all(c % 16 == 0 and c <= 128 for c in self.ch) and dim in (2, 3) and 0 < self.nout <= OUT_CH
        s = seed * 1000
        self.enc, self.dec = [], [None] * max(L - 1, 0)
        for l in range(L):
            cin, cin_s = (1, IN_CH) if l == 0 else (self.ch[l - 1], self.ch[l - 1])
            self.enc.append((ConvLayer(cin, self.ch[l], cin_s, self.ch[l], s + 2 * l, ntap),
                             ConvLayer(self.ch[l], self.ch[l], self.ch[l], self.ch[l], s + 2 * l + 1, ntap)))
        for l in range(L - 2, -1, -1):
            cin = self.ch[l + 1] + self.ch[l]
            self.dec[l] = (ConvLayer(cin, self.ch[l], cin, self.ch[l], s + 100 + 2 * l, ntap),
                           ConvLayer(self.ch[l], self.ch[l], self.ch[l], self.ch[l], s + 101 + 2 * l, ntap))
        self.out = ConvLayer(self.ch[0], self.nout, self.ch[0], OUT_CH, s + 500, ntap)
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
    def buffers(self, B, H, W, Z=1):
        key = (B, Z, H, W)
        if key in self._buf:
            return self._buf[key]
        torch = _torch()
        L = len(self.ch)
        f = 1 << (L - 1)
        assert H % f == 0 and W % f == 0 and (self.dim == 2 or Z % f == 0), "grid not divisible by the pooling depth"
        assert self.dim == 3 or Z == 1

        def t(P, C):
            return torch.zeros(P * C, dtype=torch.bfloat16, device=device())

        lv = []
        for l in range(L):
            zl = (Z >> l) if self.dim == 3 else 1
            lv.append((B * zl * (H >> l) * (W >> l), H >> l, W >> l, zl))
        bf = {"lv": lv, "X0": t(lv[0][0], IN_CH),
              "pred": torch.zeros(lv[0][0] * self.nout, dtype=torch.float32, device=device()),
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
        wf = "gf_conv3x3x3_wgrad_work" if self.dim == 3 else "gf_conv3x3_wgrad_work"
        for c in self.layers():
            lvl = self._level_of(c)
            ns, nf = ctypes.c_int(0), ctypes.c_int64(0)
            _call(wf, _i64(lv[lvl][0]), _i32(c.cout_s), _i32(c.cin_s), ctypes.byref(ns), ctypes.byref(nf))
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
        P, H, W, Z = bf["lv"][l]
        work = bf["work"]
        cg = layer.cout_s
        tail = (_dp(g), _i32(cg), _i32(layer.cout), _i32(layer.cin), _dp(work), _i64(work.numel()), _dp(layer.dw),
                _dp(layer.db), _stream())
        if self.dim == 3:
            _call("gf_conv3x3x3_wgrad", _dp(x1), _i32(c1), _i32(up1), _dp(x2), _i32(c2), _i64(P), _i32(Z), _i32(H),
                  _i32(W), *tail)
        else:
            _call("gf_conv3x3_wgrad", _dp(x1), _i32(c1), _i32(up1), _dp(x2), _i32(c2), _i64(P), _i32(H), _i32(W),
                  *tail)

    def _pool(self, kind, src, B, fine, C, dst, act=None, accum=0):
        """kind 'avg' (fine -> coarse mean), 'unpool' (coarse -> fine, weight 1/2^dim, accumulate,
        mask), 'sum' (fine -> coarse block sums, mask); `fine` = (H, W, Z) of the fine grid"""
        H, W, Z = fine
        if self.dim == 3:
            if kind == "avg":
                _call("gf_avgpool2x2x2", _dp(src), _i64(B), _i32(Z), _i32(H), _i32(W), _i32(C), _dp(dst), _stream())
            elif kind == "unpool":
                _call("gf_unpool2x2x2", _dp(src), _i64(B), _i32(Z), _i32(H), _i32(W), _i32(C), _f32(0.125),
                      _i32(accum), _dp(act), _dp(dst), _stream())
            else:
                _call("gf_sumpool2x2x2", _dp(src), _i64(B), _i32(Z), _i32(H), _i32(W), _i32(C), _dp(act), _dp(dst),
                      _stream())
        else:
            if kind == "avg":
                _call("gf_avgpool2", _dp(src), _i64(B), _i32(H), _i32(W), _i32(C), _dp(dst), _stream())
            elif kind == "unpool":
                _call("gf_unpool2", _dp(src), _i64(B), _i32(H), _i32(W), _i32(C), _f32(0.25), _i32(accum), _dp(act),
                      _dp(dst), _stream())
            else:
                _call("gf_sumpool2", _dp(src), _i64(B), _i32(H), _i32(W), _i32(C), _dp(act), _dp(dst), _stream())

    # ---- forward / backward ----
    def forward(self, B, H, W, Z=1):
        """Network on the input already in buffers(B, H, W, Z)["X0"]; returns the FP32 prediction
        [B*Z*H*W*nout]."""
        bf = self.buffers(B, H, W, Z)
        L = len(self.ch)
        x, cx = bf["X0"], IN_CH
        for l in range(L):
            P, h, w, z = bf["lv"][l]
            if l > 0:
                fine = bf["lv"][l - 1]
                self._pool("avg", bf["E"][l - 1], B, (fine[1], fine[2], fine[3]), self.ch[l - 1], bf["I"][l])
                x, cx = bf["I"][l], self.ch[l - 1]
            c1, c2 = self.enc[l]
            conv(x, cx, P, h, w, c1, bf["A"][l], flags=CF_RELU, Z=z)
            conv(bf["A"][l], self.ch[l], P, h, w, c2, bf["E"][l], flags=CF_RELU, Z=z)
        d, cd = bf["E"][L - 1], self.ch[L - 1]
        for l in range(L - 2, -1, -1):
            P, h, w, z = bf["lv"][l]
            c1, c2 = self.dec[l]
            conv(d, cd, P, h, w, c1, bf["Bd"][l], flags=CF_RELU, up1=1, x2=bf["E"][l], c2=self.ch[l], Z=z)
            conv(bf["Bd"][l], self.ch[l], P, h, w, c2, bf["D"][l], flags=CF_RELU, Z=z)
            d, cd = bf["D"][l], self.ch[l]
        P, h, w, z = bf["lv"][0]
        conv(d, cd, P, h, w, self.out, None, y32=bf["pred"], Z=z)
        return bf["pred"]

    def backward(self, B, H, W, Z=1):
        """Gradients of all layers from bf["gY"] (dL/dpred in channels 0..nout-1)."""
        bf = self.buffers(B, H, W, Z)
        L = len(self.ch)
        P0, h0, w0, z0 = bf["lv"][0]
        last = bf["D"][0] if L > 1 else bf["E"][0]
        self._wgrad(self.out, bf, 0, last, self.ch[0], bf["gY"])
        g_last = bf["gD"][0] if L > 1 else bf["gE"][0]
        conv(bf["gY"], OUT_CH, P0, h0, w0, self.out, g_last, mask=last, backward=True, Z=z0)
        for l in range(L - 1):
            P, h, w, z = bf["lv"][l]
            c1, c2 = self.dec[l]
            self._wgrad(c2, bf, l, bf["Bd"][l], self.ch[l], bf["gD"][l])
            conv(bf["gD"][l], self.ch[l], P, h, w, c2, bf["gB"][l], mask=bf["Bd"][l], backward=True, Z=z)
            dn = bf["D"][l + 1] if l + 1 < L - 1 else bf["E"][L - 1]
            self._wgrad(c1, bf, l, dn, self.ch[l + 1], bf["gB"][l], up1=1, x2=bf["E"][l], c2=self.ch[l])
            conv(bf["gB"][l], self.ch[l], P, h, w, c1, bf["gUp"][l], ysplit=self.ch[l + 1], y2=bf["gE"][l],
                 backward=True, Z=z)
            gdn = bf["gD"][l + 1] if l + 1 < L - 1 else bf["gE"][L - 1]
            self._pool("sum", bf["gUp"][l], B, (h, w, z), self.ch[l + 1], gdn, act=dn)
        for l in range(L - 1, -1, -1):
            P, h, w, z = bf["lv"][l]
            if l < L - 1:
                # skip gradient from the decoder + the pooled path of level l + 1, masked by E_l > 0
                self._pool("unpool", bf["gI"][l + 1], B, (h, w, z), self.ch[l], bf["gE"][l], act=bf["E"][l], accum=1)
            c1, c2 = self.enc[l]
            self._wgrad(c2, bf, l, bf["A"][l], self.ch[l], bf["gE"][l])
            conv(bf["gE"][l], self.ch[l], P, h, w, c2, bf["gA"][l], mask=bf["A"][l], backward=True, Z=z)
            x, cx = (bf["X0"], IN_CH) if l == 0 else (bf["I"][l], self.ch[l - 1])
            self._wgrad(c1, bf, l, x, cx, bf["gA"][l])
            if l > 0:
                conv(bf["gA"][l], self.ch[l], P, h, w, c1, bf["gI"][l], backward=True, Z=z)

    def loss_grad(self, B, H, W, target, scale=1.0, Z=1, lowpass=0.0, levels=0):
        """Loss of the prediction against target [B*Z*H*W*nout] FP32; writes gY; returns the device
        loss (a 1-element tensor).  lowpass = 0: the sum of squared errors (gf_mse).  lowpass > 0:
        gf_mse_multiscale, which adds lowpass times the squared centred means of the error over cells
        2, 4, ..., 2^levels wide, each weighted by its cell size, so that error smooth over a cell
        counts (levels + 1) times and white error barely more than once."""
        bf = self.buffers(B, H, W, Z)
        P0 = bf["lv"][0][0]
        if lowpass > 0.0:
            torch = _torch()
            zz = Z if self.dim == 3 else 1
            nf = _i64()
            _call("gf_mse_multiscale_work", _i32(self.dim), _i64(B), _i32(zz), _i32(H), _i32(W), _i32(self.nout),
                  _i32(levels), _i32(OUT_CH), ctypes.byref(nf))
            if bf.get("ms_work") is None or bf["ms_work"].numel() < nf.value:
                bf["ms_work"] = torch.zeros(nf.value, dtype=torch.float32, device=device())
            _call("gf_mse_multiscale", _i32(self.dim), _i64(B), _i32(zz), _i32(H), _i32(W), _i32(self.nout),
                  _i32(levels), _f32(lowpass), _dp(bf["pred"]), _dp(target), _f32(scale), _i32(OUT_CH), _dp(bf["gY"]),
                  _dp(bf["ms_work"]), _i64(bf["ms_work"].numel()), _dp(bf["loss"]), _stream())
            return bf["loss"]
        _call("gf_mse", _i64(P0), _i32(self.nout), _dp(bf["pred"]), _dp(target), _f32(scale), _i32(OUT_CH),
              _dp(bf["gY"]), _dp(bf["loss_part"]), _dp(bf["loss"]), _stream())
        return bf["loss"]

    def adam(self, lr):
        self.step_count += 1
        for c in self.layers():
            c.adam(lr, self.step_count)

    def set_input(self, B, H, W, G_list, shift, scale, Z=1):
        """Normalised log G of B element fields (device FP64 [Z*H*W] each) into the input buffer."""
        bf = self.buffers(B, H, W, Z)
        V = Z * H * W
        X0 = bf["X0"]
        for b, G in enumerate(G_list):
            xb = X0[b * V * IN_CH:(b + 1) * V * IN_CH]
            _call("gf_input_logG", _i64(V), _dp(G), _f64(shift), _f64(scale), _i32(IN_CH), _dp(xb), _stream())

    # ---- persistence ----
    def state(self):
        d = {"channels": np.array(self.ch, dtype=np.int64), "dim": np.array(self.dim), "nout": np.array(self.nout)}
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
            dim = int(f["dim"]) if "dim" in f.files else 2
            nout = int(f["nout"]) if "nout" in f.files else 2
            net = cls(tuple(int(c) for c in f["channels"]), dim=dim, nout=nout)
            for i, c in enumerate(net.layers()):
                c.w.copy_(torch.from_numpy(f[f"w{i}"]))
                c.b.copy_(torch.from_numpy(f[f"b{i}"]))
                c.pack()
            net.meta = {k[5:]: f[k][()] if f[k].ndim == 0 else f[k] for k in f.files if k.startswith("meta_")}
        return net


class PeriodicWarmStart:
    """Couples a WarmStartNet to the tied PCG of a periodic Regular cell, Quad4 n x n (dim 2) or
    Hex8 n^3 (dim 3), under a shear increment gamma of the control DOFs: simple shear ctrl(0, 1)
    (docs/examples/statics/periodic_elastoplastic.py:141) or pure shear ctrl(0, 1) = ctrl(1, 0).

    The independent nodes (i, j[, k]) < n are the lower corner nodes of the elements with the same
    indices, so the nodal grid and the element grid coincide.  Unknown x_u of node (i, j[, k]),
    component c: gamma * (u_fluct(c) + affine(c)) with affine u_x = y (and u_y = x for pure shear);
    the origin node is prescribed."""

    def __init__(self, pcg, tyings, n, net, sigma=0.4, target_scale=None, dim=2, shear="simple"):
        torch = _torch()
        self.P, self.ty, self.n, self.net, self.sigma, self.dim, self.shear = pcg, tyings, n, net, sigma, dim, shear
        dofs = np.asarray(tyings.dofs)
        N1 = n + 1
        r = np.arange(n)
        if dim == 2:
            J, I = np.meshgrid(r, r, indexing="ij")
            nodes = (J * N1 + I).ravel()
            coords = [I.ravel(), J.ravel()]
        else:
            K, J, I = np.meshgrid(r, r, r, indexing="ij")
            nodes = ((K * N1 + J) * N1 + I).ravel()
            coords = [I.ravel(), J.ravel(), K.ravel()]
        d = dofs[nodes][:, :dim]
        idx = np.where(d < tyings.nnu, d, -1).astype(np.int32)
        aff = np.zeros((nodes.size, dim))
        aff[:, 0] = coords[1]  # u_x = gamma * y  (ctrl(0, 1))
        if shear == "pure":
            aff[:, 1] = coords[0]  # u_y = gamma * x  (ctrl(1, 0))
        self.idx = to_device(idx.ravel(), torch.int32)
        self.affine = to_device(aff.ravel())
        self.xp_unit = np.zeros(tyings.nnp)
        self.xp_unit[tyings.control[0, 1] - tyings.nnu] = 1.0
        if shear == "pure":
            self.xp_unit[tyings.control[1, 0] - tyings.nnu] = 1.0
        self.target_scale = float(target_scale if target_scale is not None else net.meta.get("target_scale", 1.0))
        self.nv = nodes.size  # grid cells = independent nodes

    def _dims(self):
        n = self.n
        return (n, n, n) if self.dim == 3 else (n, n, 1)

    def rhs(self, gamma):
        return -self.P.apply(None, gamma * self.xp_unit)

    def predict(self, G, gamma, out=None):
        """Initial guess x_u (device FP64 [nnu]) for the modulus field G (device FP64, one per element)."""
        torch = _torch()
        H, W, Z = self._dims()
        self.net.set_input(1, H, W, [G], 0.0, 1.0 / self.sigma, Z=Z)
        pred = self.net.forward(1, H, W, Z=Z)
        x = out if out is not None else torch.zeros(self.ty.nnu, dtype=torch.float64, device=device())
        _call("gf_grid_to_dofs", _i64(self.nv * self.dim), _i32(self.dim), _dp(pred), _dp(self.idx), _dp(self.affine),
              _f64(self.target_scale), _f64(gamma), _dp(x), _stream())
        return x

    def target(self, x_u, gamma, out):
        """Centred, scaled fluctuation grid [cells * dim] FP32 of a solution x_u."""
        _call("gf_dofs_to_grid", _i64(self.nv * self.dim), _i32(self.dim), _dp(x_u), _dp(self.idx), _dp(self.affine),
              _f64(self.target_scale), _f64(gamma), _i32(1), _dp(out), _stream())
        return out

    def make_dataset(self, K, seeds, tol=1e-6, progress=None, field=None):
        """PCG solutions for the modulus fields field(seed) (default exp(sigma z_seed)); returns
        (G [ns, cells] FP64, targets [ns, cells * dim] FP32, total iterations, seconds)."""
        torch = _torch()
        ns = len(seeds)
        Gs = torch.zeros((ns, self.nv), dtype=torch.float64, device=device())
        T = torch.zeros((ns, self.nv * self.dim), dtype=torch.float32, device=device())
        Kd = to_device(K)
        its = 0
        t0 = time.perf_counter()
        for k, s in enumerate(seeds):
            g = field(s) if field is not None else np.exp(self.sigma * randn(s, self.nv))
            Gs[k] = to_device(g)
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
        H, W, Z = self._dims()
        errs = []
        for k in range(Gs.shape[0]):
            self.net.set_input(1, H, W, [Gs[k]], 0.0, 1.0 / self.sigma, Z=Z)
            p = self.net.forward(1, H, W, Z=Z).reshape(self.nv, self.dim)
            p = p - p.mean(dim=0, keepdim=True)
            t = T[k].reshape(self.nv, self.dim)
            errs.append(float(((p - t).norm() / t.norm()).item()))
        return float(np.mean(errs))

    def train(self, Gs, T, epochs, batch=4, lr=2e-3, lr_min=1e-5, log_every=1, time_limit=None, lowpass=0.0,
              levels=0):
        """Adam with cosine learning-rate decay on (Gs, T); returns per-epoch mean losses.  lowpass,
        levels: the multi-scale loss of WarmStartNet.loss_grad."""
        torch = _torch()
        H, W, Z = self._dims()
        ns = Gs.shape[0]
        nb = ns // batch
        total = epochs * nb
        rng = np.random.default_rng(11)
        net = self.net
        net.buffers(batch, H, W, Z)
        tgt = torch.zeros(batch * self.nv * self.dim, dtype=torch.float32, device=device())
        hist = []
        t0 = time.perf_counter()
        step = 0
        for ep in range(epochs):
            perm = rng.permutation(ns)
            acc = torch.zeros(1, dtype=torch.float64, device=device())
            for bi in range(nb):
                ids = perm[bi * batch:(bi + 1) * batch]
                net.set_input(batch, H, W, [Gs[i] for i in ids], 0.0, 1.0 / self.sigma, Z=Z)
                tgt.copy_(T[torch.from_numpy(ids).to(device())].reshape(-1))
                net.forward(batch, H, W, Z=Z)
                loss = net.loss_grad(batch, H, W, tgt, Z=Z, lowpass=lowpass, levels=levels)
                acc += loss.double()
                net.backward(batch, H, W, Z=Z)
                lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * step / max(total - 1, 1)))
                net.adam(float(lr_t))
                step += 1
            mean = float(acc.item()) / (nb * batch * self.nv * self.dim)
            hist.append(mean)
            if log_every and (ep % log_every == 0 or ep == epochs - 1):
                print(f"epoch {ep}: mse {mean:.4e}  {time.perf_counter() - t0:.1f} s", flush=True)
            if time_limit and time.perf_counter() - t0 > time_limit:
                break
        return hist


class PeriodicShearWarmStart(PeriodicWarmStart):
    """The 2-D simple-shear coupling (periodic Quad4 n x n, ctrl(0, 1))."""

    def __init__(self, pcg, tyings, n, net, sigma=0.4, target_scale=None):
        super().__init__(pcg, tyings, n, net, sigma=sigma, target_scale=target_scale, dim=2, shear="simple")


def splitmix_uniform(seed, n):
    """n uniform [0, 1) numbers from SplitMix64 (53-bit mantissa), identical on every platform."""
    out = np.empty(n)
    x = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def stiff_spheres(n, radius=4.0, fraction=0.2, seed=105):
    """Element mask (element (i, j, k) at (k n + j) n + i) of periodic spheres in an n^3 cell: centres
    uniform (seeded), added one at a time until the fraction of element centres inside a sphere
    reaches `fraction`; returns (mask, number of spheres)."""
    inside = np.zeros((n, n, n), dtype=bool)  # [i, j, k]
    u = splitmix_uniform(seed, 3 * 100000)
    r = int(np.ceil(radius)) + 1
    off = np.arange(-r, r + 1)
    count, total, k = 0, n ** 3, 0
    while count < fraction * total:
        cen = n * u[3 * k: 3 * k + 3]
        k += 1
        idx = [np.floor(cen[a]).astype(int) + off for a in range(3)]
        d2 = np.zeros((len(off),) * 3)
        for a in range(3):
            sh = [1, 1, 1]
            sh[a] = -1
            d2 = d2 + ((idx[a] + 0.5 - cen[a]) ** 2).reshape(sh)
        sel = d2 <= radius * radius
        I, J, K = np.meshgrid(idx[0] % n, idx[1] % n, idx[2] % n, indexing="ij")
        new = sel & ~inside[I, J, K]
        inside[I[new], J[new], K[new]] = True
        count += int(new.sum())
    return inside.transpose(2, 1, 0).ravel(), k


def sphere_field(n, seed, sigma=0.4, sphere_seed=None):
    """G = exp(sigma z) (z from `seed`), times 10 inside the stiff spheres of `sphere_seed`."""
    G = np.exp(sigma * randn(seed, n ** 3))
    mask, _ = stiff_spheres(n, seed=sphere_seed if sphere_seed is not None else seed + 1000003)
    G[mask] *= 10.0
    return G


def periodic_problem(n, sigma=0.4, seed=4, K=10.0, dim=2, spheres=False, sphere_seed=105):
    """Periodic Regular cell, Quad4 n x n plane strain (dim 2) or Hex8 n^3 (dim 3), with
    Tyings::Control + Tyings::Periodic as in docs/examples/statics/periodic_elastoplastic.py:28-47
    (iip = control DOFs + origin DOFs); per-element G = exp(sigma z), times 10 inside stiff spheres
    (dim 3, spheres=True); K constant.  Returns a namespace with the tied PCG."""
    from types import SimpleNamespace

    import goosefem_amd as gf

    mesh = gf.Mesh.Hex8.Regular(n, n, n) if dim == 3 else gf.Mesh.Quad4.Regular(n, n)
    ctrl = gf.Tyings.Control(mesh.coor, mesh.dofs)
    iip = np.concatenate((ctrl.controlDofs.ravel(), ctrl.dofs[mesh.nodesOrigin]))
    ty = gf.Tyings.Periodic(ctrl.coor, ctrl.dofs, ctrl.controlDofs, mesh.nodesPeriodic, iip)
    vec = gf.Vector(mesh.conn, ty.dofs)
    x = gf.Vector(mesh.conn, mesh.dofs).AsElement(mesh.coor)
    quad = gf.Element.Hex8.Quadrature(x) if dim == 3 else gf.Element.Quad4.QuadraturePlanar(x)
    nsph = 0
    if dim == 3 and spheres:
        G = np.exp(sigma * randn(seed, mesh.nelem))
        mask, nsph = stiff_spheres(n, seed=sphere_seed)
        G[mask] *= 10.0
    else:
        G = np.exp(sigma * randn(seed, mesh.nelem))
    Kv = np.full(mesh.nelem, K)
    P = gf.TyingsPCG(vec, quad, Kv, G, ty)
    return SimpleNamespace(mesh=mesh, ty=ty, vec=vec, quad=quad, G=G, K=Kv, P=P, n=n, sigma=sigma, dim=dim,
                           nspheres=nsph)


def train_main(argv=None):
    """Generate PCG solutions for random modulus fields, train one or more networks on the device,
    report their error on held-out fields, save the weights (npz) with the measured data-generation
    and training times."""
    import argparse
    import os

    torch = _torch()
    ap = argparse.ArgumentParser(description="train the learned PCG warm start (periodic Quad4 n x n, simple shear)")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--dim", type=int, default=2, choices=(2, 3),
                    help="2: Quad4 n^2, simple shear; 3: Hex8 n^3 with stiff spheres, pure shear")
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
    ap.add_argument("--init", default=None,
                    help="continue training these weights (one configuration; their target scale is kept)")
    ap.add_argument("--lowpass", type=float, default=0.0,
                    help="weight of the multi-scale loss term (gf_mse_multiscale): error smooth over 2^s-wide "
                         "cells, s = 1..--lowpass-levels, counts this much more per level")
    ap.add_argument("--lowpass-levels", type=int, default=0,
                    help="levels of the multi-scale term (default: the deepest the grid allows, at most 8)")
    ap.add_argument("--lr-min", type=float, default=1e-5, help="final learning rate of the cosine decay")
    a = ap.parse_args(argv)
    pb = periodic_problem(a.n, dim=a.dim, spheres=a.dim == 3)
    shear = "pure" if a.dim == 3 else "simple"
    field = (lambda sd: sphere_field(a.n, sd, pb.sigma)) if a.dim == 3 else None
    cfgs = [tuple(int(c) for c in cfg.split(",")) for cfg in a.channels.split(";")]
    lrs = [float(x) for x in a.lr.split(";")]
    lrs = lrs + [lrs[-1]] * (len(cfgs) - len(lrs))
    probe = PeriodicWarmStart(pb.P, pb.ty, a.n, WarmStartNet(cfgs[0], dim=a.dim), sigma=pb.sigma,
                              target_scale=1.0, dim=a.dim, shear=shear)
    seeds = list(range(a.seed0, a.seed0 + a.samples + a.holdout))
    Gs, T, its, t_data = probe.make_dataset(pb.K, seeds, tol=a.data_tol, progress=50 if a.dim == 2 else 10,
                                            field=field)
    init = WarmStartNet.load(a.init) if a.init else None
    if init is not None:
        cfgs = [tuple(init.ch)]
        scale = float(init.meta["target_scale"])
    else:
        scale = float(T[: a.samples].std().item())
    T.div_(scale)
    print(f"data: {len(seeds)} solves, {its} CG iterations, {t_data:.1f} s; target scale {scale:.4e}", flush=True)
    Gtr, Ttr, Gho, Tho = Gs[: a.samples], T[: a.samples], Gs[a.samples:], T[a.samples:]
    for cfg, lr in zip(cfgs, lrs):
        net = init if init is not None else WarmStartNet(cfg, dim=a.dim)
        prev_epochs = int(net.meta.get("epochs", 0)) if init is not None else 0
        prev_seconds = float(net.meta.get("train_seconds", 0.0)) if init is not None else 0.0
        W = PeriodicWarmStart(pb.P, pb.ty, a.n, net, sigma=pb.sigma, target_scale=scale, dim=a.dim, shear=shear)
        t0 = time.perf_counter()
        levels = a.lowpass_levels
        if a.lowpass > 0.0 and levels <= 0:
            levels = 0
            while levels < 8 and a.n % (2 << levels) == 0:
                levels += 1
        hist = W.train(Gtr, Ttr, a.epochs, batch=a.batch, lr=lr, lr_min=a.lr_min, time_limit=a.time_limit,
                       lowpass=a.lowpass, levels=levels)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        err = W.holdout_error(Gho, Tho) if a.holdout else float("nan")
        net.meta = {"target_scale": scale, "sigma": pb.sigma, "n": a.n, "dim": a.dim, "shear": shear,
                    "samples": a.samples, "epochs": prev_epochs + len(hist),
                    "batch": a.batch, "lr": lr, "data_seconds": t_data, "train_seconds": prev_seconds + t_train,
                    "final_mse": hist[-1],
                    "holdout_rel_error": err, "data_tol": a.data_tol, "seed0": a.seed0, "params": net.nparams(),
                    "lowpass": a.lowpass, "lowpass_levels": levels,
                    "lowpass_epochs": (int(net.meta.get("lowpass_epochs", 0)) if init is not None else 0)
                    + (len(hist) if a.lowpass > 0.0 else 0)}
        out = a.out
        if len(cfgs) > 1:
            stem, ext = os.path.splitext(a.out)
            out = f"{stem}_{'-'.join(map(str, cfg))}{ext}"
        net.save(out)
        print(f"channels {cfg} lr {lr}: {net.nparams()} parameters, {len(hist)} epochs in {t_train:.1f} s; final mse "
              f"{hist[-1]:.4e}; held-out relative error {err:.4f}; saved {out}", flush=True)


if __name__ == "__main__":
    train_main()
