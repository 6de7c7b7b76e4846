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
                                           (112, 48, (2, 4, 8))])
def test_conv3_forward(device, cin, cout, dims):
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
