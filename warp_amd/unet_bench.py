"""Throughput of the U-Net warm start's MFMA convolutions.

    python -m warp_amd.unet_bench [--dims D H W] [--reps N]

Times the network's forward pass, and forward + backward (the loss
gradient, then dX and dW/db of every layer), on a D x H x W node grid
(default 256 x 128 x 128: the nodes 0..n-1 of the 128x128x256 mesh of the
learned-warm-start configuration), with HIP events on the device's current
stream, and prints one JSON line with the algorithmic FLOPs
(2 * 27 * C_in * C_out per voxel for each 3x3x3 convolution, the same again
for dW and for dX in the backward pass, which skips the first layer's dX), the achieved TFLOP/s and their
fraction of the 2.5 PFLOP/s dense BF16 MFMA peak of the MI355X.
"""
import argparse
import json

import torch

from . import _native as N
from .unet import UNet3D

BF16_PEAK_TFLOPS = 2500.0


def conv_flops(net, dims):
    """Forward FLOPs of every 3x3x3 convolution at its level's size."""
    D, H, W = dims
    total = 0
    for lv, (ca, cb) in enumerate(net.enc):
        f = 8 ** lv
        nv = D * H * W // f
        total += 2 * 27 * (ca.cin * ca.cout + cb.cin * cb.cout) * nv
    for lv, (ca, cb) in enumerate(net.dec):
        f = 8 ** (2 - lv)
        nv = D * H * W // f
        total += 2 * 27 * (ca.cin * ca.cout + cb.cin * cb.cout) * nv
    total += 2 * 27 * net.out.cin * net.out.cout * D * H * W
    return total


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--dims", type=int, nargs=3, default=[256, 128, 128])
    p.add_argument("--reps", type=int, default=5)
    a = p.parse_args()
    dims = tuple(a.dims)
    if any(d % 8 for d in dims):
        raise SystemExit("each of D, H, W must be divisible by 8")
    dev = torch.device("cuda", 0)
    net = UNet3D(dev, seed=1)
    # give the zero-initialised output layer weights so the backward pass
    # runs through every layer with nonzero gradients
    g = torch.Generator().manual_seed(2)
    net.out.w.copy_(torch.randn(net.out.w.shape, generator=g).to(dev) * 0.05)
    net.out.pack()
    nv = dims[0] * dims[1] * dims[2]
    x = (torch.randn(nv * 16, generator=g) * 0.5).to(torch.bfloat16).view(torch.int16).to(dev)
    tgt = torch.randn(nv, generator=g).to(torch.float32).to(dev)
    gout = torch.empty(nv * 16, dtype=torch.int16, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    st = N.stream_of(dev)
    s = torch.cuda.current_stream(dev)

    def fwd():
        return net.forward(x, dims)

    def fwd_bwd():
        net.zero_grad()
        out = net.forward(x, dims, keep=True)
        loss.zero_()
        N.call("wm_unet_mse", out.data_ptr(), tgt.data_ptr(), nv, gout.data_ptr(),
               loss.data_ptr(), st)
        net.backward(gout)

    def timed(f):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(a.reps):
            f()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / a.reps

    ms_f = timed(fwd)
    ms_fb = timed(fwd_bwd)
    flops_f = conv_flops(net, dims)
    # backward: dW of every layer, dX of every layer but the first
    first = 2 * 27 * net.enc[0][0].cin * net.enc[0][0].cout * nv
    flops_fb = flops_f + flops_f + (flops_f - first)
    tf_f = flops_f / (ms_f * 1e-3) / 1e12
    tf_fb = flops_fb / (ms_fb * 1e-3) / 1e12
    print(json.dumps({"dims": list(dims), "reps": a.reps,
                      "forward_ms": ms_f, "forward_tflop": flops_f / 1e12,
                      "forward_tflops": tf_f, "forward_frac": tf_f / BF16_PEAK_TFLOPS,
                      "fwd_bwd_ms": ms_fb, "fwd_bwd_tflop": flops_fb / 1e12,
                      "fwd_bwd_tflops": tf_fb, "fwd_bwd_frac": tf_fb / BF16_PEAK_TFLOPS,
                      "peak_tflops": BF16_PEAK_TFLOPS}), flush=True)


if __name__ == "__main__":
    main()
