"""The learned initial guess of the multigrid solve.

No counterpart in the reference, whose solve starts from the previous
step's phi (scripts/fieldsolver.py:1653-1658).  Here the guess is the
linear extrapolation in time, 2 phi_(n-1) - phi_(n-2), plus a correction
the 3-D U-Net (warp_amd.unet) predicts for it.  The extrapolation's error
is L^-1 (rho_n - 2 rho_(n-1) + rho_(n-2)) with the conductors held at zero
(the voltages cancel), so the network sees that second time difference of
the source (and the source itself and the conductor fraction) and predicts
the error.  It works at half resolution: a coarse voxel is a 2x2x2 block of
mesh nodes; its output is interpolated back to the nodes.  Inputs are
normalised by their largest magnitude and the output is scaled back by the
same factor times cphys (a constant fixed at training), so the prediction
is linear in the source's scale.

The multigrid still runs to mgtol from the guess, so every solve keeps the
reference's stopping rule; and the guess falls back to the extrapolation
alone (the plain method) whenever the network's residual is not smaller.
All of the arithmetic runs in libwarpmi kernels (wm_unet_*); the weights
are a numpy .npz file (no pickles).
"""
import numpy as np
import torch

from . import _native as N
from .globals import eps0
from .unet import CHANNELS, SLOPE, UNet3D, conductor_mask


def default_weights():
    """The committed network trained on the 128x128x256 problem."""
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_c3.npz")


def half_dims(solver):
    nx, ny, nz = solver.nx, solver.ny, solver.nz
    if nx % 16 or ny % 16 or nz % 16:
        raise ValueError("the learned guess needs nx, ny, nz divisible by 16 (got %d, %d, %d)"
                         % (nx, ny, nz))
    return (nz // 2, ny // 2, nx // 2)


def save_weights(path, net, cphys, meta=None):
    """Layer weights and biases (fp32) plus cphys and metadata, as .npz."""
    arrs = {"cphys": np.array(cphys), "channels": np.array(net.channels),
            "slope": np.array(net.slope)}
    for i, c in enumerate(net.layers()):
        arrs["w%02d" % i] = c.w.cpu().numpy()
        arrs["b%02d" % i] = c.b.cpu().numpy()
    for k, v in (meta or {}).items():
        arrs["meta_" + k] = np.array(v)
    np.savez(path, **arrs)


def load_weights(path, device):
    """(UNet3D, cphys, metadata dict) from a file written by save_weights."""
    with np.load(path, allow_pickle=False) as f:
        net = UNet3D(device, channels=tuple(int(c) for c in f["channels"]),
                     slope=float(f["slope"]))
        for i, c in enumerate(net.layers()):
            c.w.copy_(torch.from_numpy(f["w%02d" % i]))
            c.b.copy_(torch.from_numpy(f["b%02d" % i]))
            c.pack()
        meta = {k[5:]: f[k] for k in f.files if k.startswith("meta_")}
        return net, float(f["cphys"]), meta


class LearnedGuess:
    """Per-solver state of the learned guess: the coarse sources of the last
    three solves, the conductor masks and the network."""

    def __init__(self, solver, net=None, cphys=1.0):
        self.solver = solver
        self.dev = solver.device
        self.dims = half_dims(solver)
        self.nh = self.dims[0] * self.dims[1] * self.dims[2]
        self.net = net if net is not None else UNet3D(self.dev)
        self.cphys = float(cphys)
        self.ring = []             # coarse sources, newest first
        self.d2 = torch.empty(self.nh, dtype=torch.float32, device=self.dev)
        self.mx = torch.zeros(2, dtype=torch.int32, device=self.dev)
        self.x = torch.empty(self.nh * 16, dtype=torch.int16, device=self.dev)
        self.mask = None
        self.maskh = None
        self._mask_key = None
        self.applied = 0
        self.fallbacks = 0

    def _masks(self):
        key = getattr(self.solver, "_cond_gen", None)
        if self.mask is None or key != self._mask_key:
            cond = conductor_mask(self.solver)
            nz, ny, nx = self.solver.nz, self.solver.ny, self.solver.nx
            m = cond.reshape(nz, ny, nx).to(torch.float64)
            # coarse conductor fraction, through the same block mean (a node
            # array with the nz+1, ny+1, nx+1 layout the kernel expects)
            full = torch.zeros((nz + 1, ny + 1, nx + 1), dtype=torch.float64, device=self.dev)
            full[:nz, :ny, :nx] = m
            self.maskh = torch.empty(self.nh, dtype=torch.float32, device=self.dev)
            N.call("wm_unet_half_mean", full.data_ptr(), nx, ny, nz, 0, self.maskh.data_ptr(),
                   N.stream_of(self.dev))
            # the correction goes only to the region the beam lives in
            self.mask = torch.from_numpy(outside_main_region(cond.cpu().numpy().reshape(
                nz, ny, nx)).reshape(-1)).to(self.dev)
            self._mask_key = key
        return self.mask, self.maskh

    def observe(self, rho):
        """Adds the coarse mean of rho (the source of the coming solve)."""
        s = self.solver
        r = torch.empty(self.nh, dtype=torch.float32, device=self.dev)
        N.call("wm_unet_half_mean", rho.data_ptr(), s.nx, s.ny, s.nz, 0, r.data_ptr(),
               N.stream_of(self.dev))
        self.ring = [r] + self.ring[:2]

    def ready(self):
        return len(self.ring) == 3

    def features(self):
        """Second difference and maxima of the last three coarse sources,
        and the network input built from them."""
        s = N.stream_of(self.dev)
        _, maskh = self._masks()
        self.mx.zero_()
        r0, r1, r2 = self.ring
        N.call("wm_unet_second_diff", r0.data_ptr(), r1.data_ptr(), r2.data_ptr(), self.nh,
               self.d2.data_ptr(), self.mx.data_ptr(), s)
        N.call("wm_unet_learn_input", self.d2.data_ptr(), r0.data_ptr(), maskh.data_ptr(),
               self.mx.data_ptr(), self.nh, self.x.data_ptr(), s)
        return self.x

    def correct(self, phi):
        """phi += the network's estimate of the extrapolation error."""
        mask, _ = self._masks()
        x = self.features()
        out = self.net.forward(x, self.dims)
        s = self.solver
        N.call("wm_unet_learn_apply", out.data_ptr(), self.mx.data_ptr(), self.cphys, s.nx, s.ny,
               s.nz, mask.data_ptr(), phi.data_ptr(), N.stream_of(self.dev))


def outside_main_region(cond):
    """uint8 (nz, ny, nx): 1 where the learned correction must not go --
    conductor interior nodes (cond = 1) and every free node that is not
    6-connected through free nodes to the centre of the mesh (pockets
    enclosed by a conductor's interior shell, e.g. the corners outside a
    pipe).  The multigrid converges slowly on errors in such pockets, whose
    potential the extrapolation already has exactly; the network's small
    output there only costs V-cycles.  When the centre is a conductor node
    the largest free region is kept."""
    from scipy import ndimage
    free = cond == 0
    lab, n = ndimage.label(free)
    if n == 0:
        return np.ones_like(cond, dtype=np.uint8)
    nz, ny, nx = cond.shape
    c = lab[nz // 2, ny // 2, nx // 2]
    if c == 0:
        c = int(np.argmax(np.bincount(lab.ravel())[1:])) + 1
    return (lab != c).astype(np.uint8)


def residual_max(solver, phi):
    """max |residual3d| of phi for the solver's rho (conductors included),
    through the solver's level-0 entry point."""
    res = torch.zeros_like(phi)
    reps0c = solver.mgparam / (eps0 * 2. * (1. / solver.dx ** 2 + 1. / solver.dy ** 2 +
                                             1. / solver.dz ** 2))
    scaled = (solver.rho * reps0c).contiguous()
    N.call("wm_mg3d_residual", solver._h, 0, phi.data_ptr(), scaled.data_ptr(), res.data_ptr(),
           solver.mgparam, solver.mgform, 0, int(solver.lcndbndy), solver.icndbndy,
           N.stream_of(solver.device))
    return float(res.abs().max())
