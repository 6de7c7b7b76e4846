"""Learned warm start for the multigrid solve: a small 3-D U-Net that maps
(rho, conductor mask, previous-step phi) to the change of phi, trained on
pairs the multigrid itself generates.  The network only supplies the
initial guess; the multigrid then runs to the usual mgtol, so every solve
keeps the reference's stopping rule, and plain multigrid (the previous phi
as the guess) is the fallback.

Architecture: 4 levels with 16/32/48/64 channels, two 3x3x3 convolutions
(leaky ReLU) per level on the way down and up, 2x2x2 average pooling,
nearest-neighbour upsampling with skip concatenation, and a final 3x3x3
convolution to one channel.  bf16 activations and weights, fp32
accumulation and fp32 master weights trained with Adam.  Every forward and
backward operation is a libwarpmi kernel (wm_unet_*); torch only holds the
device memory.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import _native as N

CHANNELS = (16, 32, 48, 64)
SLOPE = 0.1


class _Conv:
    """One 3x3x3 convolution: fp32 master weights w[co][ci][27], bias b,
    their gradients and Adam moments, and the packed bf16 copies."""

    def __init__(self, cin, cout, device, rng, act=True):
        self.cin, self.cout, self.act = cin, cout, act
        std = math.sqrt(2.0 / (27 * cin))
        w = rng.standard_normal((cout, cin, 27)).astype(np.float32) * std
        self.w = torch.from_numpy(w).to(device)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device)
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)
        self.m = [torch.zeros_like(self.w), torch.zeros_like(self.w),
                  torch.zeros_like(self.b), torch.zeros_like(self.b)]
        n = int(N.lib().wm_unet_pack_size(cin, cout))
        self.wp = torch.empty(n, dtype=torch.int16, device=device)
        self.wt = torch.empty(n, dtype=torch.int16, device=device)
        self.pack()

    def pack(self):
        N.call("wm_unet_pack", self.w.data_ptr(), self.cin, self.cout, self.wp.data_ptr(),
               self.wt.data_ptr(), N.stream_of(self.w.device))

    def params(self):
        return [(self.w, self.gw, self.m[0], self.m[1]), (self.b, self.gb, self.m[2], self.m[3])]


class UNet3D:
    """The warm-start network on a D x H x W node grid (each divisible
    by 8)."""

    def __init__(self, device, seed=0, channels=CHANNELS, slope=SLOPE):
        self.device = torch.device(device)
        self.slope = slope
        rng = np.random.default_rng(seed)
        c0, c1, c2, c3 = channels
        mk = lambda a, b, act=True: _Conv(a, b, self.device, rng, act)
        self.enc = [[mk(16, c0), mk(c0, c0)], [mk(c0, c1), mk(c1, c1)],
                    [mk(c1, c2), mk(c2, c2)], [mk(c2, c3), mk(c3, c3)]]
        self.dec = [[mk(c3 + c2, c2), mk(c2, c2)], [mk(c2 + c1, c1), mk(c1, c1)],
                    [mk(c1 + c0, c0), mk(c0, c0)]]
        self.out = mk(c0, 16, act=False)  # channel 0 is the output
        # the last layer starts at zero: the untrained network predicts no
        # change, i.e. the previous phi, which is the plain method
        self.out.w.zero_()
        self.out.pack()
        self.channels = channels
        self.step_count = 0

    def layers(self):
        return [c for lv in self.enc for c in lv] + [c for lv in self.dec for c in lv] + [self.out]

    # -- helpers --------------------------------------------------------
    def _t(self, nvox, c):
        return torch.empty(nvox * c, dtype=torch.int16, device=self.device)

    def _conv(self, conv, x, dims, y=None, yf=None):
        D, H, W = dims
        N.call("wm_unet_conv3", x.data_ptr(), conv.cin, y.data_ptr() if y is not None else None,
               yf.data_ptr() if yf is not None else None, conv.cout, D, H, W,
               conv.wp.data_ptr(), conv.b.data_ptr(), self.slope if conv.act else 1.0,
               N.stream_of(self.device))

    # -- forward --------------------------------------------------------
    def forward(self, x, dims, keep=False):
        """x: bf16 (int16 bits) input of 16 channels on dims (D, H, W).
        Returns the fp32 16-channel output (channel 0 is the prediction);
        keep=True stores what backward needs."""
        D, H, W = dims
        s = N.stream_of(self.device)
        saved = []
        skips = []
        h = x
        d = (D, H, W)
        for l, (ca, cb) in enumerate(self.enc):
            nv = d[0] * d[1] * d[2]
            if l > 0:
                p = self._t(nv, ca.cin)
                N.call("wm_unet_pool", h.data_ptr(), 2 * d[0], 2 * d[1], 2 * d[2], ca.cin,
                       p.data_ptr(), 0, s)
                h = p
            a = self._t(nv, ca.cout)
            self._conv(ca, h, d, y=a)
            b = self._t(nv, cb.cout)
            self._conv(cb, a, d, y=b)
            saved.append((ca, h, a, d))
            saved.append((cb, a, b, d))
            skips.append((b, d))
            h = b
            if l < 3:
                d = (d[0] // 2, d[1] // 2, d[2] // 2)
        # decoder
        for l, (ca, cb) in enumerate(self.dec):
            sk, ds = skips[2 - l]
            nv = ds[0] * ds[1] * ds[2]
            cup = ca.cin - sk.numel() // nv
            cat = self._t(nv, ca.cin)
            N.call("wm_unet_upcat", h.data_ptr(), d[0], d[1], d[2], cup, sk.data_ptr(),
                   ca.cin - cup, cat.data_ptr(), s)
            d = ds
            a = self._t(nv, ca.cout)
            self._conv(ca, cat, d, y=a)
            b = self._t(nv, cb.cout)
            self._conv(cb, a, d, y=b)
            saved.append((ca, cat, a, d))
            saved.append((cb, a, b, d))
            h = b
        nv = D * H * W
        out = torch.empty(nv * 16, dtype=torch.float32, device=self.device)
        self._conv(self.out, h, (D, H, W), yf=out)
        saved.append((self.out, h, None, (D, H, W)))
        if keep:
            self._saved = saved
        return out

    # -- backward -------------------------------------------------------
    def backward(self, gout):
        """gout: bf16 gradient of the 16-channel output.  Accumulates the
        weight and bias gradients of every layer."""
        s = N.stream_of(self.device)
        saved = self._saved
        k = len(saved) - 1

        def conv_back(entry, g, need_dx=True):
            conv, xin, y, d = entry
            nv = d[0] * d[1] * d[2]
            if conv.act:
                N.call("wm_unet_act_bwd", g.data_ptr(), y.data_ptr(), self.slope, nv * conv.cout, s)
            N.call("wm_unet_conv3_wgrad", xin.data_ptr(), conv.cin, g.data_ptr(), conv.cout, d[0],
                   d[1], d[2], conv.gw.data_ptr(), conv.gb.data_ptr(), s)
            if not need_dx:
                return None
            gx = self._t(nv, conv.cin)
            N.call("wm_unet_conv3", g.data_ptr(), conv.cout, gx.data_ptr(), None, conv.cin, d[0],
                   d[1], d[2], conv.wt.data_ptr(), None, 1.0, s)
            return gx

        g = conv_back(saved[k], gout)
        k -= 1
        skip_grads = []
        for l in range(3):
            g = conv_back(saved[k], g); k -= 1           # second conv of the level
            gcat = conv_back(saved[k], g); k -= 1        # first conv, on the concat
            ca, cat, a, d = saved[k + 1]
            nv = d[0] * d[1] * d[2]
            cskip = self.channels[l]  # backward visits decoder levels 0, 1, 2
            cup = ca.cin - cskip
            gs = self._t(nv, cskip)
            N.call("wm_unet_chcopy", gcat.data_ptr(), ca.cin, cup, gs.data_ptr(), cskip, 0, cskip,
                   nv, s)
            skip_grads.append(gs)
            dc = (d[0] // 2, d[1] // 2, d[2] // 2)
            gu = self._t(dc[0] * dc[1] * dc[2], cup)
            N.call("wm_unet_up", gcat.data_ptr(), dc[0], dc[1], dc[2], cup, gu.data_ptr(), ca.cin, 1,
                   s)
            g = gu
        # encoder, bottom level first
        for l in range(3, -1, -1):
            if l < 3:
                gs = skip_grads[l]
                N.call("wm_unet_add", g.data_ptr(), gs.data_ptr(), g.numel(), s)
            g = conv_back(saved[k], g); k -= 1
            g = conv_back(saved[k], g, need_dx=l > 0); k -= 1
            if l > 0:
                ca, xin, a, d = saved[k + 1]
                full = (2 * d[0], 2 * d[1], 2 * d[2])
                gp = self._t(full[0] * full[1] * full[2], ca.cin)
                N.call("wm_unet_pool", g.data_ptr(), full[0], full[1], full[2], ca.cin,
                       gp.data_ptr(), 1, s)
                g = gp
        self._saved = None

    def zero_grad(self):
        for c in self.layers():
            c.gw.zero_()
            c.gb.zero_()

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        s = N.stream_of(self.device)
        for c in self.layers():
            for w, g, m1, m2 in c.params():
                N.call("wm_unet_adam", w.data_ptr(), g.data_ptr(), m1.data_ptr(), m2.data_ptr(),
                       w.numel(), lr, b1, b2, eps, self.step_count, s)
            c.pack()


def make_input(solver, phi_prev, mask, srho, sphi):
    """Network input on the solver's nodes 0..n-1 (bf16, 16 channels)."""
    nx, ny, nz = solver.nx, solver.ny, solver.nz
    x = torch.empty(nx * ny * nz * 16, dtype=torch.int16, device=solver.device)
    N.call("wm_unet_input", solver.rho.data_ptr(), phi_prev.data_ptr(),
           mask.data_ptr() if mask is not None else None, nx, ny, nz, srho, sphi, x.data_ptr(),
           N.stream_of(solver.device))
    return x


def conductor_mask(solver):
    """1 on interior conductor nodes (0..n-1), from the installed lists."""
    nx, ny, nz = solver.nx, solver.ny, solver.nz
    m = np.zeros((nz, ny, nx), dtype=np.uint8)
    if solver.hasconductors():
        L = solver.getconductorobject().lists()
        e = L["int_istart"][1] if len(L["int_istart"]) > 1 else 0
        idx = np.asarray(L["int_indx"][:e], dtype=np.int64).reshape(-1, 3)
        k = (idx[:, 0] < nx) & (idx[:, 1] < ny) & (idx[:, 2] < nz)
        m[idx[k, 2], idx[k, 1], idx[k, 0]] = 1
    return torch.from_numpy(m.reshape(-1)).to(solver.device)
