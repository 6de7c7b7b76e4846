"""Training data, training loop and evaluation of the U-Net warm start.

Pairs come from the multigrid itself: for a seeded random beam (Gaussian
or uniform ellipse, centroid offsets, current) the solver computes phi for
the beam at z0 (the previous step) and at z0 + vz*dt (this step); the
network learns phi - phi_prev from (rho, conductor mask, phi_prev).  The
sources are the beams' charge densities evaluated on the mesh nodes.
"""
import math

import numpy as np
import torch

from . import _native as N
from .globals import eps0
from .unet import UNet3D, conductor_mask, make_input


def beam_rho(solver, rng, zshift=0.0, params=None):
    """Charge density of a random beam on the solver's nodes, and the
    parameters drawn (re-used with zshift for the next step)."""
    if params is None:
        kind = "gauss" if rng.random() < 0.5 else "kv"
        params = dict(kind=kind, a=rng.uniform(0.002, 0.006), b=rng.uniform(0.002, 0.006),
                      x0=rng.uniform(-0.003, 0.003), y0=rng.uniform(-0.003, 0.003),
                      z0=rng.uniform(0.025, 0.055), sz=rng.uniform(0.008, 0.015),
                      tilt=rng.uniform(-0.02, 0.02), current=10 ** rng.uniform(-4, -2))
    p = params
    dev = solver.device
    z = solver.zmmin + solver.dz * torch.arange(solver.nz + 1, device=dev, dtype=torch.float64)
    y = solver.ymmin + solver.dy * torch.arange(solver.ny + 1, device=dev, dtype=torch.float64)
    x = solver.xmmin + solver.dx * torch.arange(solver.nx + 1, device=dev, dtype=torch.float64)
    Z, Y, X = torch.meshgrid(z, y, x, indexing="ij")
    zc = p["z0"] + zshift
    xs = X - p["x0"] - p["tilt"] * (Z - zc)
    ys = Y - p["y0"]
    lam = p["current"] / 1.0e6  # line charge at vz = 1e6 m/s, C/m
    if p["kind"] == "gauss":
        tr = torch.exp(-0.5 * (xs / p["a"]) ** 2 - 0.5 * (ys / p["b"]) ** 2) / (2 * math.pi * p["a"] * p["b"])
    else:
        tr = ((xs / p["a"]) ** 2 + (ys / p["b"]) ** 2 <= 1.).to(torch.float64) / (math.pi * p["a"] * p["b"])
    lz = torch.exp(-0.5 * ((Z - zc) / p["sz"]) ** 2)
    return lam * tr * lz, params


def generate_pairs(solver, n, seed, dz_step=1.0e-3, mgtol=1e-9):
    """n training pairs: (rho, phi_prev, phi) FP64 tensors on the GPU."""
    rng = np.random.default_rng(seed)
    out = []
    save_tol = solver.mgtol
    solver.mgtol = mgtol
    for _ in range(n):
        rho0, p = beam_rho(solver, rng)
        rho1, _ = beam_rho(solver, rng, zshift=dz_step, params=p)
        solver.phi = 0.
        solver.rho = rho0
        solver.solve()
        phi0 = solver.phi.clone()
        solver.rho = rho1
        solver.solve()
        out.append((rho1, phi0, solver.phi.clone()))
    solver.mgtol = save_tol
    return out


class WarmStart:
    """A trained network plus the scales of its inputs and output."""

    def __init__(self, solver, seed=0):
        self.solver = solver
        self.net = UNet3D(solver.device, seed=seed)
        self.mask = conductor_mask(solver)
        self.dims = (solver.nz, solver.ny, solver.nx)
        self.srho = self.sphi = self.sdphi = 1.0

    def fit_scales(self, pairs):
        self.srho = 1.0 / max(float(r.abs().max()) for r, _, _ in pairs)
        self.sphi = 1.0 / max(float(p0.abs().max()) for _, p0, _ in pairs)
        self.sdphi = 1.0 / max(float((p1 - p0).abs().max()) for _, p0, p1 in pairs)

    def _target(self, phi_prev, phi):
        d = (phi - phi_prev)[1:-2, 1:-2, 1:-2] * self.sdphi
        return d.reshape(-1).to(torch.float32).contiguous()

    def _input(self, rho, phi_prev):
        self.solver.rho = rho
        return make_input(self.solver, phi_prev, self.mask, self.srho, self.sphi)

    def train(self, pairs, steps, lr=1e-3, seed=0):
        """Adam on the mean-square error of the scaled change of phi;
        returns the loss history (one float per step)."""
        rng = np.random.default_rng(seed)
        net, dev = self.net, self.solver.device
        inputs = [(self._input(r, p0), self._target(p0, p1)) for r, p0, p1 in pairs]
        nv = self.dims[0] * self.dims[1] * self.dims[2]
        g = torch.empty(nv * 16, dtype=torch.int16, device=dev)
        loss = torch.zeros(1, dtype=torch.float32, device=dev)
        hist = []
        s = N.stream_of(dev)
        for _ in range(steps):
            x, tgt = inputs[int(rng.integers(len(inputs)))]
            net.zero_grad()
            out = net.forward(x, self.dims, keep=True)
            loss.zero_()
            N.call("wm_unet_mse", out.data_ptr(), tgt.data_ptr(), nv, g.data_ptr(),
                   loss.data_ptr(), s)
            net.backward(g)
            net.adam_step(lr)
            hist.append(float(loss.item()))
        return hist

    def predict(self, rho, phi_prev):
        """phi_prev plus the predicted change at the unpinned nodes."""
        x = self._input(rho, phi_prev)
        out = self.net.forward(x, self.dims)
        phi = phi_prev.clone()
        N.call("wm_unet_apply", out.data_ptr(), self.mask.data_ptr(), self.solver.nx,
               self.solver.ny, self.solver.nz, self.sdphi, phi.data_ptr(),
               N.stream_of(self.solver.device))
        return phi


def residual_max(solver, rho, phi):
    """max |residual| of phi for the source rho on the finest level
    (residual3d, as the V-cycle computes it; conductors included)."""
    res = torch.zeros_like(phi)
    reps0c = solver.mgparam / (eps0 * 2. * (1. / solver.dx ** 2 + 1. / solver.dy ** 2 +
                                             1. / solver.dz ** 2))
    scaled = rho * reps0c
    N.call("wm_mg3d_residual", solver._h, 0, phi.data_ptr(), scaled.contiguous().data_ptr(),
           res.data_ptr(), solver.mgparam, solver.mgform, 0, int(solver.lcndbndy),
           solver.icndbndy, N.stream_of(solver.device))
    return float(res.abs().max())


def choose_guess(solver, ws, rho, phi_prev):
    """The network's guess unless its residual is larger than that of the
    previous phi (then the plain method's guess: a fallback)."""
    guess = ws.predict(rho, phi_prev)
    if residual_max(solver, rho, guess) < residual_max(solver, rho, phi_prev):
        return guess, False
    return phi_prev, True


def vcycles_from(solver, rho, phi0, mgtol=1e-6):
    """V-cycles of one solve to mgtol from the initial guess phi0."""
    save = solver.mgtol
    solver.mgtol = mgtol
    solver.rho = rho
    solver.phi = phi0
    solver.solve()
    solver.mgtol = save
    return solver.mgiters, solver.mgerror
