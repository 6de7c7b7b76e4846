"""Particle storage and seeded loaders.

:class:`ParticleGroup` mirrors the reference's ParticleGroup (top.v
Particles group): structure-of-arrays FP64 columns xp, yp, zp, uxp, uyp,
uzp, gaminv, ex, ey, ez, and per-species sq, sm, sw, ins, nps, here as
torch tensors resident on the GPU.  Loads follow addparticles
(scripts/particles.py:1407-1646): particles outside the mesh are culled.
"""
import math

import numpy as np
import torch

from .globals import amu, echarge


class ParticleGroup:
    COLUMNS = ("xp", "yp", "zp", "uxp", "uyp", "uzp", "gaminv", "ex", "ey", "ez")

    def __init__(self, device="cuda", dtype=torch.float64):
        self.device = torch.device(device)
        self.dtype = dtype
        for c in self.COLUMNS:
            setattr(self, c, torch.zeros(0, dtype=dtype, device=self.device))
        self.sq, self.sm, self.sw, self.dtscale = [], [], [], []
        self.ins, self.nps = [], []  # 1-based start, count (as the reference)
        self.ndts = []

    @property
    def ns(self):
        return len(self.sq)

    @property
    def npmax(self):
        return int(self.xp.numel())

    def addspecies(self, charge, mass, weight):
        self.sq.append(float(charge))
        self.sm.append(float(mass))
        self.sw.append(float(weight))
        self.dtscale.append(1.)
        self.ndts.append(1)
        self.ins.append(self.npmax + 1)
        self.nps.append(0)
        return self.ns - 1

    def addparticles(self, js, x, y, z, ux, uy, uz, gaminv=None):
        """Append particles to species js (which must be the last species
        that received particles, as the arrays are stored contiguously)."""
        x = torch.as_tensor(x, dtype=self.dtype, device=self.device)
        n = int(x.numel())
        cols = dict(xp=x, yp=y, zp=z, uxp=ux, uyp=uy, uzp=uz)
        if gaminv is None:
            gaminv = torch.ones(n, dtype=self.dtype, device=self.device)
        cols["gaminv"] = gaminv
        for c in ("ex", "ey", "ez"):
            cols[c] = torch.zeros(n, dtype=self.dtype, device=self.device)
        assert self.ins[js] - 1 + self.nps[js] == self.npmax, \
            "particles must be appended to the last species"
        for c in self.COLUMNS:
            v = torch.as_tensor(cols[c], dtype=self.dtype, device=self.device).reshape(-1)
            setattr(self, c, torch.cat([getattr(self, c), v]))
        self.nps[js] += n

    def species_slice(self, js):
        i = self.ins[js] - 1
        return slice(i, i + self.nps[js])

    def compact(self, keep):
        """Drop particles where keep is False, species by species
        (the reference's particle boundary handling removes lost
        particles)."""
        new_ins, new_nps, parts = [], [], []
        start = 0
        for js in range(self.ns):
            sl = self.species_slice(js)
            k = keep[sl]
            idx = torch.nonzero(k).reshape(-1) + sl.start
            parts.append(idx)
            new_ins.append(start + 1)
            new_nps.append(int(idx.numel()))
            start += int(idx.numel())
        idx = torch.cat(parts) if parts else torch.zeros(0, dtype=torch.long, device=self.device)
        for c in self.COLUMNS:
            setattr(self, c, getattr(self, c)[idx].contiguous())
        self.ins, self.nps = new_ins, new_nps


def particlesortxyzwithcopy(pgroup, dx, dy, dz, xmmin, ymmin, zmmin, nx, ny, nz):
    """Orders each species' particles by mesh block (top.F:5095-5203):
    block int(|x-xmmin|/dx) + (nx+1)*(iy + (ny+1)*iz), stable within a
    block.  Every column is permuted (particlecopysorteddata).  On the GPU:
    cell keys + radix sort + one gather pass over the columns."""
    import ctypes as C
    from . import _native as N
    if pgroup.npmax == 0:
        return
    if pgroup.device.type != "cuda":
        raise RuntimeError("particlesortxyzwithcopy runs on the HIP device")
    s = N.stream_of(pgroup.device)
    new = {c: torch.empty_like(getattr(pgroup, c)) for c in pgroup.COLUMNS}
    k = len(pgroup.COLUMNS)
    covered = 0
    for js in range(pgroup.ns):
        sl = pgroup.species_slice(js)
        n = sl.stop - sl.start
        if n == 0:
            continue
        covered += n
        need = int(N.lib().wm_particle_sort_worksize(n, nx, ny, nz))
        work = torch.empty(need, dtype=torch.uint8, device=pgroup.device)
        perm = torch.empty(n, dtype=torch.int32, device=pgroup.device)
        N.call("wm_particle_sort", n, pgroup.xp[sl].data_ptr(), pgroup.yp[sl].data_ptr(),
               pgroup.zp[sl].data_ptr(), xmmin, ymmin, zmmin, dx, dy, dz, nx, ny, nz,
               perm.data_ptr(), work.data_ptr(), s)
        src = (C.c_void_p * k)(*[getattr(pgroup, c)[sl].data_ptr() for c in pgroup.COLUMNS])
        dst = (C.c_void_p * k)(*[new[c][sl].data_ptr() for c in pgroup.COLUMNS])
        N.call("wm_permute_columns", n, k, src, dst, perm.data_ptr(), s)
        del work, perm
    if covered != pgroup.npmax:
        raise RuntimeError("particle arrays hold particles outside every species")
    for c in pgroup.COLUMNS:
        setattr(pgroup, c, new[c])


def gaussian_weight(current, sigmaz, vz, charge, npart):
    """Macroparticle weight for a beam of peak current I whose line charge is
    a Gaussian of rms length sigmaz: sw = I*sqrt(2 pi)*sigmaz/(vz*q*N)
    (the reference's sw = |I*zlen/(v*q*N)| with the peak line charge,
    pywarp90/w3d.F:7142-7143, 7162-7172)."""
    return current * math.sqrt(2 * math.pi) * sigmaz / (vz * charge * npart)


def uniform_weight(current, length, vz, charge, npart):
    """sw = I*zlen/(vz*q*N) for a beam uniform over length
    (pywarp90/w3d.F:7142-7143)."""
    return current * length / (vz * charge * npart)


def _cull(x, y, z, ux, uy, uz, bounds):
    xmmin, xmmax, ymmin, ymmax, zmmin, zmmax = bounds
    keep = ((x >= xmmin) & (x < xmmax) & (y >= ymmin) & (y < ymmax) &
            (z >= zmmin) & (z < zmmax))
    return x[keep], y[keep], z[keep], ux[keep], uy[keep], uz[keep]


def load_gaussian(npart, sigma, zmean, vz, vspread, bounds, seed, charge=echarge,
                  mass=39 * amu, current=1e-3, device="cuda"):
    """Seeded Gaussian bunch: sigma = (sx, sy, sz), mean (0, 0, zmean),
    velocity (0, 0, vz) plus a Gaussian spread of vspread*vz per component.
    Returns (ParticleGroup, species index, number loaded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.normal(0., sigma[0], npart)
    y = rng.normal(0., sigma[1], npart)
    z = zmean + rng.normal(0., sigma[2], npart)
    ux = rng.normal(0., vspread * vz, npart)
    uy = rng.normal(0., vspread * vz, npart)
    uz = vz + rng.normal(0., vspread * vz, npart)
    x, y, z, ux, uy, uz = _cull(x, y, z, ux, uy, uz, bounds)
    pg = ParticleGroup(device)
    js = pg.addspecies(charge, mass, gaussian_weight(current, sigma[2], vz, charge, npart))
    pg.addparticles(js, x, y, z, ux, uy, uz)
    return pg, js, len(x)


def load_kv_ellipse(npart, a, b, zmin, zmax, vz, vspread, bounds, seed, charge=echarge,
                    mass=39 * amu, current=1e-3, device="cuda"):
    """Seeded KV-like beam: uniform in the ellipse (x/a)^2+(y/b)^2 <= 1 and
    uniform in z over [zmin, zmax]; velocities as load_gaussian."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = np.sqrt(rng.random(npart))
    th = 2 * np.pi * rng.random(npart)
    x = a * r * np.cos(th)
    y = b * r * np.sin(th)
    z = zmin + (zmax - zmin) * rng.random(npart)
    ux = rng.normal(0., vspread * vz, npart)
    uy = rng.normal(0., vspread * vz, npart)
    uz = vz + rng.normal(0., vspread * vz, npart)
    x, y, z, ux, uy, uz = _cull(x, y, z, ux, uy, uz, bounds)
    pg = ParticleGroup(device)
    js = pg.addspecies(charge, mass, uniform_weight(current, zmax - zmin, vz, charge, npart))
    pg.addparticles(js, x, y, z, ux, uy, uz)
    return pg, js, len(x)
