"""Training of the learned initial guess (warp_amd.learned) on data the
multigrid itself generates.

Data: time-stepped runs of randomised beams (particles.load_c3_beam) on the
128x128x256 mesh with the pipe and the annulus electrode at 0-20 kV, run by
the step driver with the plain method (the previous phi as the guess, the
reference's).  After every step the coarse source and phi are recorded;
from the third step on each step gives one sample: the network input (the
second time difference of the coarse source, the coarse source, the
conductor fraction) and the target, the error of the linear extrapolation
2 phi_(n-1) - phi_(n-2) averaged over coarse voxels, in units of
cphys * max|d2|.  Training: Adam on the mean-square error, one sample per
step, through the library's forward and backward kernels.

    python -m warp_amd.unet_train --out warp_amd/data/warmstart_c3.npz \
        --trajectories 800 --max-samples 12000 --train-steps 20000
"""
import argparse
import math
import sys
import time

import numpy as np
import torch

from . import _native as N
from .learned import LearnedGuess, save_weights
from .unet import UNet3D

VOLTAGES = (0., 5.0e3, 10.0e3, 15.0e3, 20.0e3)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_solver(voltage, device, n=(128, 128, 256), **kw):
    """The 128x128x256 problem's solver with the annulus at `voltage`."""
    import warp_amd
    from . import configs
    cfg = dict(configs.CONFIGS["c1"], n=n)
    s = configs.make_solver(cfg, device=device, mgverbose=-1, **kw)
    s.uninstallconductor(s.conductordatalist[1][0])
    ann = warp_amd.ZAnnulus(0.012, 0.018, 0.005, voltage, zcent=0.0325, condid=2)
    s.installconductor(ann)
    return s


class Recorder:
    """After-step hook: keeps the last three phi and coarse sources and
    appends (d2, r0, mx, target) samples."""

    def __init__(self, solver, samples, max_samples):
        self.s = solver
        self.lg = LearnedGuess(solver)
        self.phis = []
        self.samples = samples
        self.max_samples = max_samples
        self.tgt_scale = None

    def __call__(self):
        s, lg = self.s, self.lg
        lg.observe(s.rho)
        self.phis = [s.phi.clone()] + self.phis[:2]
        if not lg.ready() or len(self.samples) >= self.max_samples:
            return
        lg.features()
        tgt = torch.empty(lg.nh, dtype=torch.float32, device=lg.dev)
        p0, p1, p2 = self.phis
        # target in units of max|d2| (cphys = 1); rescaled once cphys is known
        N.call("wm_unet_learn_target", p0.data_ptr(), p1.data_ptr(), p2.data_ptr(), s.nx, s.ny,
               s.nz, lg.mx.data_ptr(), 1.0, tgt.data_ptr(), N.stream_of(lg.dev))
        self.samples.append((lg.d2.clone(), lg.ring[0].clone(), lg.mx.clone(), tgt))


def collect(device, ntraj, nsteps, npart, seed, max_samples, solvers=None, log_every=10):
    """Samples from ntraj randomised runs of nsteps steps each."""
    from . import registersolver, unregistersolver, stepper, configs
    from .globals import top
    from .particles import load_c3_beam
    rng = np.random.default_rng(seed)
    solvers = solvers if solvers is not None else {}
    samples = []
    t0 = time.time()
    for k in range(ntraj):
        v = VOLTAGES[int(rng.integers(len(VOLTAGES)))]
        if v not in solvers:
            solvers[v] = make_solver(v, device)
        s = solvers[v]
        pg, p = load_c3_beam(rng, npart, configs.DOMAIN, configs.VZ, configs.VSPREAD,
                             device=device)
        saved = {a: getattr(top, a) for a in ("pgroup", "dt", "it", "time", "nhist",
                                              "lspecial", "particlesortinterval", "pbound0",
                                              "pboundnz", "pboundxy")}
        rec = Recorder(s, samples, max_samples)
        try:
            top.pgroup, top.dt, top.it, top.time, top.nhist = pg, configs.DT, 0, 0., 0
            top.particlesortinterval = 0
            top.pbound0 = top.pboundnz = top.pboundxy = 0
            s.phi = 0.
            registersolver(s)
            s.sortparticles(pg)
            stepper.generate()
            rec()
            stepper.installafterstep(rec)
            stepper.step(nsteps)
        finally:
            if rec in stepper._afterstep:
                stepper.uninstallafterstep(rec)
            unregistersolver(s)
            for a, val in saved.items():
                setattr(top, a, val)
        if (k + 1) % log_every == 0 or k + 1 == ntraj:
            log("collect: %d/%d runs, %d samples, %.0f s (last: %s, %.1f kV, %.2f mA, "
                "V-cycles %s)" % (k + 1, ntraj, len(samples), time.time() - t0,
                                  p.get("mix", p["kind"]), v / 1e3, p["current"] * 1e3,
                                  s.mgiters_history[-3:]))
        if len(samples) >= max_samples:
            break
    return samples, solvers


def fit_cphys(samples):
    """cphys: the median of max|target| over the samples, so that the
    network's targets are of order one; the stored targets are rescaled."""
    m = [float(t.abs().max()) for _, _, _, t in samples]
    c = float(np.median(m))
    for _, _, _, t in samples:
        t.div_(c)
    return c


def train(lg, samples, steps, lr=3e-4, lr_end=3e-5, seed=0, holdout=None, log_every=250):
    """Adam with a cosine learning-rate decay; returns the loss history and
    the held-out ratio of the network's mean-square error to the
    extrapolation's (predicting zero)."""
    rng = np.random.default_rng(seed)
    net, dev = lg.net, lg.dev
    _, maskh = lg._masks()
    s = N.stream_of(dev)
    g = torch.empty(lg.nh * 16, dtype=torch.int16, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    x = torch.empty(lg.nh * 16, dtype=torch.int16, device=dev)
    hist = []
    t0 = time.time()

    def inp(sample):
        d2, r0, mx, _ = sample
        N.call("wm_unet_learn_input", d2.data_ptr(), r0.data_ptr(), maskh.data_ptr(),
               mx.data_ptr(), lg.nh, x.data_ptr(), s)
        return x

    def evaluate(sset):
        num = den = 0.
        for smp in sset:
            out = net.forward(inp(smp), lg.dims)
            loss.zero_()
            N.call("wm_unet_mse", out.data_ptr(), smp[3].data_ptr(), lg.nh, g.data_ptr(),
                   loss.data_ptr(), s)
            num += float(loss.item())
            den += float((smp[3] ** 2).mean())
        return num / den

    for k in range(steps):
        net.step_lr = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * k / steps))
        smp = samples[int(rng.integers(len(samples)))]
        net.zero_grad()
        out = net.forward(inp(smp), lg.dims, keep=True)
        loss.zero_()
        N.call("wm_unet_mse", out.data_ptr(), smp[3].data_ptr(), lg.nh, g.data_ptr(),
               loss.data_ptr(), s)
        net.backward(g)
        net.adam_step(net.step_lr)
        hist.append(float(loss.item()) / max(float((smp[3] ** 2).mean()), 1e-30))
        if (k + 1) % log_every == 0 or k + 1 == steps:
            msg = "train: step %d/%d, lr %.2e, relative loss (last %d) %.4f, %.0f s" % (
                k + 1, steps, net.step_lr, log_every, np.mean(hist[-log_every:]),
                time.time() - t0)
            if holdout:
                msg += ", held-out %.4f" % evaluate(holdout)
            log(msg)
    return hist, (evaluate(holdout) if holdout else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--trajectories", type=int, default=120)
    ap.add_argument("--steps-per-run", type=int, default=24)
    ap.add_argument("--particles", type=int, default=2_000_000)
    ap.add_argument("--max-samples", type=int, default=3000)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=2026)
    a = ap.parse_args()
    dev = torch.device("cuda", 0)
    samples, solvers = collect(dev, a.trajectories, a.steps_per_run, a.particles, a.seed,
                               a.max_samples)
    hold, _ = collect(dev, 4, a.steps_per_run, a.particles, a.seed + 1, 10 ** 9, solvers)
    cphys = fit_cphys(samples)
    for _, _, _, t in hold:
        t.div_(cphys)
    log("cphys %.4e from %d samples (%d held out)" % (cphys, len(samples), len(hold)))
    lg = LearnedGuess(next(iter(solvers.values())), UNet3D(dev, seed=a.seed), cphys)
    hist, ratio = train(lg, samples, a.train_steps, lr=a.lr, seed=a.seed, holdout=hold[::4])
    # runs used: collection stops once max_samples are in (each run gives
    # steps_per_run - 1 samples, from its third solve on)
    nruns = min(a.trajectories, -(-len(samples) // max(1, a.steps_per_run - 1)))
    save_weights(a.out, lg.net, cphys,
                 meta=dict(samples=len(samples), runs=nruns,
                           solves=nruns * (a.steps_per_run + 1), train_steps=a.train_steps,
                           heldout_ratio=ratio, mesh=[128, 128, 256]))
    log("saved %s: held-out error / extrapolation error (mean square) %.4f" % (a.out, ratio))


if __name__ == "__main__":
    main()
