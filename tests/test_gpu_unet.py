"""The U-Net warm start's kernels against torch's CPU convolution and
autograd in FP64 (on the same bf16-rounded operands), and the trained
network's effect on the multigrid solve."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


def _bf16(a):
    """Round an FP32/FP64 tensor to bf16 and back (CPU, FP64)."""
    return a.to(torch.bfloat16).to(torch.float64)


def _to_dev_bits(a_cl, device):
    """channels-last bf16 bits (int16) on the device from a CPU tensor."""
    return a_cl.contiguous().to(torch.bfloat16).view(torch.int16).reshape(-1).to(device)


def _from_dev_bits(t, shape):
    return t.cpu().view(torch.bfloat16).to(torch.float64).reshape(shape)


def _ref_conv(x, w, b, slope):
    """x (C,D,H,W), w (co, ci, 27) -> leaky_relu(conv3d(x, w, pad 1) + b)."""
    co, ci = w.shape[:2]
    y = F.conv3d(x[None], w.reshape(co, ci, 3, 3, 3), b, padding=1)[0]
    return torch.where(y > 0, y, y * slope)


@pytest.mark.parametrize("cin,cout,dims", [(16, 16, (4, 6, 8)), (32, 48, (4, 4, 4)),
                                           (112, 48, (2, 4, 8)), (16, 32, (256, 64, 16)),
                                           (16, 80, (256, 32, 16))])
def test_conv3_forward(device, cin, cout, dims):
    """Small grids run one row of 16 voxels per wavefront, the two large
    ones 4 (32 output channels) and 2 (80) rows."""
    from warp_amd import _native as N
    D, H, W = dims
    g = torch.Generator().manual_seed(cin * 7 + cout)
    x = _bf16(torch.randn(cin, D, H, W, generator=g, dtype=torch.float64))
    w = _bf16(torch.randn(cout, cin, 27, generator=g, dtype=torch.float64) * 0.1)
    b = torch.randn(cout, generator=g, dtype=torch.float64).to(torch.float32)
    wd = w.to(torch.float32).to(device)
    wp = torch.empty(int(N.lib().wm_unet_pack_size(cin, cout)), dtype=torch.int16, device=device)
    wt = torch.empty_like(wp)
    s = N.stream_of(device)
    N.call("wm_unet_pack", wd.data_ptr(), cin, cout, wp.data_ptr(), wt.data_ptr(), s)
    xd = _to_dev_bits(x.permute(1, 2, 3, 0), device)
    yf = torch.empty(D * H * W * cout, dtype=torch.float32, device=device)
    bd = b.to(device)
    N.call("wm_unet_conv3", xd.data_ptr(), cin, None, yf.data_ptr(), cout, D, H, W, wp.data_ptr(),
           bd.data_ptr(), 0.1, s)
    ref = _ref_conv(x, w, b.to(torch.float64), 0.1).permute(1, 2, 3, 0)
    got = yf.cpu().to(torch.float64).reshape(D, H, W, cout)
    assert (got - ref).abs().max() <= 2e-5 * ref.abs().max() + 1e-6


@pytest.mark.parametrize("cin,cout,dims", [(16, 16, (19, 22, 40)), (32, 16, (9, 13, 33)),
                                           (48, 16, (11, 9, 20)), (16, 16, (1, 3, 5)),
                                           (16, 32, (13, 17, 35)), (32, 32, (10, 9, 20))])
def test_conv3_zmarch_forward(device, cin, cout, dims):
    """The z-marching kernel (16 output channels with up to 48 input
    channels, 32 with up to 32): several blocks in x, y and z with partial
    tiles and chunks, both output forms; the bf16 output is the fp32 output
    rounded to nearest even."""
    from warp_amd import _native as N
    D, H, W = dims
    g = torch.Generator().manual_seed(cin + D)
    x = _bf16(torch.randn(cin, D, H, W, generator=g, dtype=torch.float64))
    w = _bf16(torch.randn(cout, cin, 27, generator=g, dtype=torch.float64) * 0.1)
    b = torch.randn(cout, generator=g, dtype=torch.float64).to(torch.float32)
    wd = w.to(torch.float32).to(device)
    wp = torch.empty(int(N.lib().wm_unet_pack_size(cin, cout)), dtype=torch.int16, device=device)
    wt = torch.empty_like(wp)
    s = N.stream_of(device)
    N.call("wm_unet_pack", wd.data_ptr(), cin, cout, wp.data_ptr(), wt.data_ptr(), s)
    xd = _to_dev_bits(x.permute(1, 2, 3, 0), device)
    yf = torch.full((D * H * W * cout,), float("nan"), dtype=torch.float32, device=device)
    yb = torch.zeros(D * H * W * cout, dtype=torch.int16, device=device)
    bd = b.to(device)
    N.call("wm_unet_conv3", xd.data_ptr(), cin, yb.data_ptr(), yf.data_ptr(), cout, D, H, W,
           wp.data_ptr(), bd.data_ptr(), 0.1, s)
    ref = _ref_conv(x, w, b.to(torch.float64), 0.1).permute(1, 2, 3, 0)
    got = yf.cpu().reshape(D, H, W, cout)
    assert torch.isfinite(got).all()
    assert (got.to(torch.float64) - ref).abs().max() <= 2e-5 * ref.abs().max() + 1e-6
    assert torch.equal(yb.cpu().view(torch.bfloat16).reshape(D, H, W, cout),
                       got.to(torch.bfloat16))


@pytest.mark.parametrize("c", [16, 12])
def test_pool_up_copy_add_act_bitwise(device, c):
    """Pooling, upsampling, channel copy, add and the activation's backward
    pass, in their 16-byte form (c = 16) and their one-value form (c = 12),
    against the same fp32 arithmetic in the same order on the CPU, bit for
    bit (bf16 rounding to nearest even)."""
    from warp_amd import _native as N
    s = N.stream_of(device)
    D, H, W = 6, 4, 10
    g = torch.Generator().manual_seed(c)
    bits = lambda t: t.to(torch.bfloat16).view(torch.int16)
    val = lambda b: b.view(torch.bfloat16).to(torch.float32)
    x = bits(torch.randn(D, H, W, c, generator=g))
    xd = x.reshape(-1).to(device)
    # average pooling and its adjoint
    y = torch.zeros(D // 2 * H // 2 * W // 2 * c, dtype=torch.int16, device=device)
    N.call("wm_unet_pool", xd.data_ptr(), D, H, W, c, y.data_ptr(), 0, s)
    xv = val(x)
    acc = torch.zeros(D // 2, H // 2, W // 2, c)
    for k in range(8):
        acc = acc + xv[(k >> 2)::2, ((k >> 1) & 1)::2, (k & 1)::2]
    assert torch.equal(y.cpu(), bits(0.125 * acc).reshape(-1))
    gb = torch.zeros(D * H * W * c, dtype=torch.int16, device=device)
    N.call("wm_unet_pool", y.data_ptr(), D, H, W, c, gb.data_ptr(), 1, s)
    up = val(y.cpu().reshape(D // 2, H // 2, W // 2, c))
    up = up.repeat_interleave(2, 0).repeat_interleave(2, 1).repeat_interleave(2, 2)
    assert torch.equal(gb.cpu(), bits(0.125 * up).reshape(-1))
    # nearest upsampling into channels [0, c) of a 2c-channel tensor, and its adjoint
    cy = 2 * c
    yu = torch.zeros(8 * D * H * W * cy, dtype=torch.int16, device=device)
    N.call("wm_unet_up", xd.data_ptr(), D, H, W, c, yu.data_ptr(), cy, 0, s)
    ref = x.repeat_interleave(2, 0).repeat_interleave(2, 1).repeat_interleave(2, 2)
    got = yu.cpu().reshape(2 * D, 2 * H, 2 * W, cy)
    assert torch.equal(got[..., :c], ref) and not got[..., c:].any()
    gu = bits(torch.randn(2 * D, 2 * H, 2 * W, cy, generator=g))
    gx = torch.zeros(D * H * W * c, dtype=torch.int16, device=device)
    N.call("wm_unet_up", gu.reshape(-1).to(device).data_ptr(), D, H, W, c, gx.data_ptr(), cy, 1, s)
    gv = val(gu)[..., :c]
    acc = torch.zeros(D, H, W, c)
    for k in range(8):
        acc = acc + gv[(k >> 2)::2, ((k >> 1) & 1)::2, (k & 1)::2]
    assert torch.equal(gx.cpu(), bits(acc).reshape(-1))
    # the decoder's concatenated input [up(x) | skip]
    sk = bits(torch.randn(2 * D, 2 * H, 2 * W, c, generator=g))
    cat = torch.zeros(8 * D * H * W * cy, dtype=torch.int16, device=device)
    N.call("wm_unet_upcat", xd.data_ptr(), D, H, W, c, sk.reshape(-1).to(device).data_ptr(), c,
           cat.data_ptr(), s)
    assert torch.equal(cat.cpu().reshape(2 * D, 2 * H, 2 * W, cy), torch.cat([ref, sk], dim=3))
    # channel copy (aligned offsets, then an offset of 4 channels)
    for soff, doff in ((0, c), (4, 0)):
        dst = torch.zeros(D * H * W * cy, dtype=torch.int16, device=device)
        cc = c - soff
        N.call("wm_unet_chcopy", xd.data_ptr(), c, soff, dst.data_ptr(), cy, doff, cc, D * H * W, s)
        got = dst.cpu().reshape(D, H, W, cy)
        assert torch.equal(got[..., doff:doff + cc], x[..., soff:])
        assert got.count_nonzero() == (x[..., soff:] != 0).sum()
    # a += b and the leaky ReLU's backward pass
    b = bits(torch.randn(D, H, W, c, generator=g))
    a = xd.clone()
    N.call("wm_unet_add", a.data_ptr(), b.reshape(-1).to(device).data_ptr(), a.numel(), s)
    assert torch.equal(a.cpu(), bits(val(x) + val(b)).reshape(-1))
    gr = a.clone()
    N.call("wm_unet_act_bwd", gr.data_ptr(), xd.data_ptr(), 0.1, gr.numel(), s)
    av = val(a.cpu())
    want = torch.where(val(x.reshape(-1)) > 0, a.cpu(), bits(av * torch.tensor(0.1, dtype=torch.float32)))
    assert torch.equal(gr.cpu(), want)


def test_conv3_backward(device):
    """dX (the mirrored convolution), dW and db against torch autograd."""
    from warp_amd import _native as N
    cin, cout, (D, H, W) = 16, 32, (4, 4, 8)
    g = torch.Generator().manual_seed(3)
    x = _bf16(torch.randn(cin, D, H, W, generator=g, dtype=torch.float64)).requires_grad_()
    w = _bf16(torch.randn(cout, cin, 27, generator=g, dtype=torch.float64) * 0.1).requires_grad_()
    b = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    gy = _bf16(torch.randn(cout, D, H, W, generator=g, dtype=torch.float64))
    y = F.conv3d(x[None], w.reshape(cout, cin, 3, 3, 3), b, padding=1)[0]
    (y * gy).sum().backward()
    s = N.stream_of(device)
    wd = w.detach().to(torch.float32).to(device)
    wp = torch.empty(int(N.lib().wm_unet_pack_size(cin, cout)), dtype=torch.int16, device=device)
    wt = torch.empty_like(wp)
    N.call("wm_unet_pack", wd.data_ptr(), cin, cout, wp.data_ptr(), wt.data_ptr(), s)
    xd = _to_dev_bits(x.detach().permute(1, 2, 3, 0), device)
    gd = _to_dev_bits(gy.permute(1, 2, 3, 0), device)
    gx = torch.empty(D * H * W * cin, dtype=torch.int16, device=device)
    N.call("wm_unet_conv3", gd.data_ptr(), cout, gx.data_ptr(), None, cin, D, H, W, wt.data_ptr(),
           None, 1.0, s)
    dw = torch.zeros(cout * cin * 27, dtype=torch.float32, device=device)
    db = torch.zeros(cout, dtype=torch.float32, device=device)
    N.call("wm_unet_conv3_wgrad", xd.data_ptr(), cin, gd.data_ptr(), cout, D, H, W, dw.data_ptr(),
           db.data_ptr(), s)
    gxr = x.grad.permute(1, 2, 3, 0)
    got = _from_dev_bits(gx, (D, H, W, cin))
    assert (got - gxr).abs().max() <= 1e-2 * gxr.abs().max()   # dX is stored in bf16
    dwg = dw.cpu().to(torch.float64).reshape(cout, cin, 27)
    assert (dwg - w.grad).abs().max() <= 1e-5 * w.grad.abs().max()
    assert (db.cpu().to(torch.float64) - b.grad).abs().max() <= 1e-5 * b.grad.abs().max()


@pytest.mark.parametrize("cin,cout,dims", [(16, 16, (8, 16, 16)), (16, 64, (4, 6, 8)),
                                           (112, 48, (2, 4, 8)), (64, 64, (4, 8, 8)),
                                           (80, 32, (3, 5, 7)), (48, 16, (1, 2, 2)),
                                           (32, 48, (16, 8, 8))])
def test_conv3_weight_gradient_shapes(device, cin, cout, dims):
    """dW and db of every channel combination the network uses (and odd
    grid sizes: partial voxel steps, rows shorter than a step) against torch
    autograd in FP64 on the same bf16 operands."""
    from warp_amd import _native as N
    D, H, W = dims
    g = torch.Generator().manual_seed(cin + 3 * cout + D)
    x = _bf16(torch.randn(cin, D, H, W, generator=g, dtype=torch.float64)).requires_grad_()
    w = torch.zeros(cout, cin, 27, dtype=torch.float64, requires_grad=True)
    b = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    gy = _bf16(torch.randn(cout, D, H, W, generator=g, dtype=torch.float64))
    y = F.conv3d(x[None], w.reshape(cout, cin, 3, 3, 3), b, padding=1)[0]
    (y * gy).sum().backward()
    s = N.stream_of(device)
    xd = _to_dev_bits(x.detach().permute(1, 2, 3, 0), device)
    gd = _to_dev_bits(gy.permute(1, 2, 3, 0), device)
    dw = torch.full((cout * cin * 27,), 0.5, dtype=torch.float32, device=device)
    db = torch.zeros(cout, dtype=torch.float32, device=device)
    N.call("wm_unet_conv3_wgrad", xd.data_ptr(), cin, gd.data_ptr(), cout, D, H, W, dw.data_ptr(),
           db.data_ptr(), s)
    dwg = dw.cpu().to(torch.float64).reshape(cout, cin, 27) - 0.5   # accumulates onto dw
    scale = w.grad.abs().max()
    assert (dwg - w.grad).abs().max() <= 1e-5 * scale + 1e-6
    assert (db.cpu().to(torch.float64) - b.grad).abs().max() <= 1e-5 * b.grad.abs().max()


class _RefUNet(torch.nn.Module):
    """The same network in torch (CPU, FP64) with the device weights."""

    def __init__(self, net):
        super().__init__()
        self.layers = [torch.nn.Parameter(_bf16(c.w.detach().cpu().to(torch.float64)))
                       for c in net.layers()]
        self.bias = [torch.nn.Parameter(c.b.detach().cpu().to(torch.float64)) for c in net.layers()]
        self.acts = [c.act for c in net.layers()]
        self.slope = net.slope

    def conv(self, i, x):
        w = self.layers[i]
        co, ci = w.shape[:2]
        y = F.conv3d(x[None], w.reshape(co, ci, 3, 3, 3), self.bias[i], padding=1)[0]
        return torch.where(y > 0, y, y * self.slope) if self.acts[i] else y

    def forward(self, x):
        skips, h, i = [], x, 0
        for lv in range(4):
            if lv > 0:
                h = F.avg_pool3d(h[None], 2)[0]
            h = self.conv(i + 1, self.conv(i, h))
            i += 2
            skips.append(h)
        for lv in range(3):
            up = F.interpolate(h[None], scale_factor=2, mode="nearest")[0]
            h = torch.cat([up, skips[2 - lv]], 0)
            h = self.conv(i + 1, self.conv(i, h))
            i += 2
        return self.conv(i, h)


def test_unet_forward_backward_against_torch(device):
    """Whole network: output and every layer's weight gradient agree with
    torch autograd in FP64 on the same bf16 weights (bf16 activations on
    the GPU: a few percent; gradient directions to cosine > 0.99)."""
    from warp_amd import _native as N
    from warp_amd.unet import UNet3D
    D, H, W = 8, 16, 16
    net = UNet3D(device, seed=4)
    g = torch.Generator().manual_seed(5)
    # the last layer starts at zero (no change predicted); give it weights
    # so that gradients reach every layer
    net.out.w.copy_(torch.randn(net.out.w.shape, generator=g) * 0.05)
    net.out.pack()
    x = _bf16(torch.randn(16, D, H, W, generator=g, dtype=torch.float64))
    x[3:] = 0
    ref = _RefUNet(net)
    yr = ref(x)
    tgt = torch.randn(D, H, W, generator=g, dtype=torch.float64)
    loss = ((yr[0] - tgt) ** 2).mean()
    loss.backward()
    xd = _to_dev_bits(x.permute(1, 2, 3, 0), device)
    out = net.forward(xd, (D, H, W), keep=True)
    o = out.cpu().to(torch.float64).reshape(D, H, W, 16)[..., 0]
    assert (o - yr[0].detach()).abs().max() <= 0.05 * yr[0].abs().max()
    gout = torch.empty(D * H * W * 16, dtype=torch.int16, device=device)
    lossd = torch.zeros(1, dtype=torch.float32, device=device)
    tg = tgt.reshape(-1).to(torch.float32).to(device)
    N.call("wm_unet_mse", out.data_ptr(), tg.data_ptr(), D * H * W, gout.data_ptr(),
           lossd.data_ptr(), N.stream_of(device))
    assert float(lossd.item()) == pytest.approx(float(loss), rel=0.05)
    net.zero_grad()
    net.backward(gout)
    for c, pr in zip(net.layers(), ref.layers):
        a = c.gw.cpu().to(torch.float64).reshape(-1)
        b = pr.grad.reshape(-1)
        cos = float((a @ b) / (a.norm() * b.norm() + 1e-300))
        assert cos > 0.99, cos


def test_unet_training_overfits_one_sample(device):
    """Adam on one fixed sample drives the loss down by 10x (at a learning
    rate inside Adam's stable range for this problem: at 3e-3 the loss
    oscillates and the outcome depends on the last bits of the gradients)."""
    from warp_amd import _native as N
    from warp_amd.unet import UNet3D
    D, H, W = 8, 16, 16
    net = UNet3D(device, seed=6)
    g = torch.Generator().manual_seed(7)
    x = torch.randn(D, H, W, 16, generator=g)
    x[..., 3:] = 0
    xd = _to_dev_bits(x, device)
    tgt = torch.sin(torch.linspace(0, 3, D * H * W)).to(torch.float32).to(device)
    gout = torch.empty(D * H * W * 16, dtype=torch.int16, device=device)
    loss = torch.zeros(1, dtype=torch.float32, device=device)
    hist = []
    for _ in range(100):
        net.zero_grad()
        out = net.forward(xd, (D, H, W), keep=True)
        loss.zero_()
        N.call("wm_unet_mse", out.data_ptr(), tgt.data_ptr(), D * H * W, gout.data_ptr(),
               loss.data_ptr(), N.stream_of(device))
        net.backward(gout)
        net.adam_step(1e-3)
        hist.append(float(loss.item()))
    assert hist[-1] < 0.1 * hist[0], hist[::10]


def test_learned_guess_trains_and_runs(device):
    """The learned-guess pipeline at 32x32x64 (pipe + annulus): samples from
    two short randomised runs of the step driver, a short training through
    the library's kernels, then the network's error of the extrapolation
    residual on held-out samples below that of predicting zero, and a
    step-driver run with guess="learned" whose every solve reaches mgtol."""
    import warp_amd
    from warp_amd import learned as LG, stepper, configs, unet_train as T
    from warp_amd.globals import top
    from warp_amd.unet import UNet3D
    n = (32, 32, 64)
    solvers = {v: T.make_solver(v, device, n=n) for v in T.VOLTAGES}
    samples, _ = T.collect(device, 4, 10, 200000, 5, 10 ** 9, solvers)
    hold, _ = T.collect(device, 1, 10, 200000, 6, 10 ** 9, solvers)
    cphys = T.fit_cphys(samples)
    for *_, t in hold:
        t.div_(cphys)
    lg = LG.LearnedGuess(solvers[0.], UNet3D(device, seed=3), cphys)
    hist, ratio = T.train(lg, samples, 300, lr=3e-4, holdout=hold, log_every=100)
    print("\nlearned guess at 32x32x64: %d samples, held-out mean-square error / "
          "extrapolation's %.3f" % (len(samples), ratio))
    assert len(samples) == 4 * 9 and ratio < 1.0
    # the step driver with the learned guess
    import os
    import tempfile
    path = os.path.join(tempfile.mkdtemp(), "w.npz")
    LG.save_weights(path, lg.net, cphys)
    s = T.make_solver(10.0e3, device, n=n, guess="learned", learned_weights=path)
    pg = configs.load_beam(configs.CONFIGS["c1"], device, npart=200000)
    saved = {a: getattr(top, a) for a in ("pgroup", "dt", "it", "time", "pbound0", "pboundnz",
                                          "particlesortinterval")}
    try:
        top.pgroup, top.dt, top.it, top.time = pg, configs.DT, 0, 0.
        top.pbound0 = top.pboundnz = 0
        top.particlesortinterval = 0
        warp_amd.registersolver(s)
        stepper.generate()
        stepper.step(8)
        st = s.learned_guess_state()
        assert s.mgerror <= s.mgtol
        assert st.applied + st.fallbacks == 7  # every solve from the third on
        print("learned guess in the step driver: applied %d, fell back %d, V-cycles %s"
              % (st.applied, st.fallbacks, s.mgiters_history))
    finally:
        warp_amd.unregistersolver(s)
        for a, v in saved.items():
            setattr(top, a, v)
