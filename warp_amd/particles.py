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
        """Length of the particle arrays (the reference's npmax): live
        particles of species js are [ins-1, ins-1+nps) of it; clearpart
        leaves the slots past a species' live block unused."""
        return int(self.xp.numel())

    @property
    def nplive(self):
        """Live particles of all species."""
        return int(sum(self.nps))

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
        end = self.ins[js] - 1 + self.nps[js]
        assert all(self.ins[k] <= self.ins[js] for k in range(self.ns)), \
            "particles must be appended to the last species"
        if end < self.npmax:
            # slots clearpart left past the species' live block
            for c in self.COLUMNS:
                setattr(self, c, getattr(self, c)[:end])
        for c in self.COLUMNS:
            v = torch.as_tensor(cols[c], dtype=self.dtype, device=self.device).reshape(-1)
            setattr(self, c, torch.cat([getattr(self, c), v]))
        self.nps[js] += n

    def species_slice(self, js):
        i = self.ins[js] - 1
        return slice(i, i + self.nps[js])

    def _columns(self, sl):
        return [getattr(self, c)[sl] for c in self.COLUMNS]

    def _lostwork(self, n):
        from . import _native as N
        need = int(N.lib().wm_clearpart_worksize(n))
        w = getattr(self, "_lpwork", None)
        if w is None or w.numel() < need:
            w = self._lpwork = torch.empty(need, dtype=torch.uint8, device=self.device)
            self._lpcount = torch.zeros(1, dtype=torch.int64, device=self.device)
        return w

    def _remove_lost(self, js, fillmethod, mesh):
        """processlostpart -> clearpart for species js on the GPU; with
        mesh = (xmmin, ymmin, zmmin, dx, dy, dz, nx, ny, nz, zgrid, zbound)
        the tests of particleboundaries3d flag the lost particles first
        (zbound 0: z absorbs, 2: z periodic).  Returns the number removed."""
        import ctypes as C
        from . import _native as N
        sl = self.species_slice(js)
        n = sl.stop - sl.start
        if fillmethod not in (1, 2):
            raise NotImplementedError("clearlostpart = %s: fill methods 1 and 2 only"
                                      % fillmethod)
        if n == 0:
            return 0
        if self.device.type != "cuda":
            raise RuntimeError("particle boundaries run on the HIP device")
        cols = self._columns(sl)
        work = self._lostwork(n)
        src = (C.c_void_p * len(cols))(*[t.data_ptr() for t in cols])
        dst, dcols = None, None
        if fillmethod == 2:
            dcols = [torch.empty_like(t) for t in cols]
            dst = (C.c_void_p * len(cols))(*[t.data_ptr() for t in dcols])
        s = N.stream_of(self.device)
        if mesh is None:
            N.call("wm_clearpart", n, len(cols), src, fillmethod, dst,
                   self._lpcount.data_ptr(), work.data_ptr(), s)
        else:
            N.call("wm_particleboundaries3d", n, len(cols), src, *mesh, fillmethod, dst,
                   self._lpcount.data_ptr(), work.data_ptr(), s)
        live = int(self._lpcount.item())
        if fillmethod == 2:
            for t, d in zip(cols, dcols):
                t[:live].copy_(d[:live])
        self.nps[js] = live
        return n - live

    def clearpart(self, js=-1, fillmethod=1):
        """clearpart (top.F:4730-4893): removes the particles flagged lost
        (gaminv <= 0) of species js (all when js < 0) with the given fill
        method (1: holes filled from the end, 2: order kept)."""
        for k in (range(self.ns) if js < 0 else [js]):
            self._remove_lost(k, fillmethod, None)


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
    for js in range(pgroup.ns):
        sl = pgroup.species_slice(js)
        n = sl.stop - sl.start
        if n == 0:
            continue
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
    # slots outside every species' live block hold nothing: left unset
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


def load_c3_beam(rng, npart, bounds, vz, vspread, charge=echarge, mass=39 * amu,
                 device="cuda", kind=None, **over):
    """A randomised beam of the learned warm start's training recipe:
    transverse Gaussian, KV-like (uniform in an ellipse), waterbag (uniform
    in a 4-D ellipsoid, whose projection has density 1 - r^2) or a mix of
    two of them; sizes a, b in [2, 6] mm, centroid offsets up to 3 mm,
    tilts up to 20 mrad (in the velocity and in the position along z), a
    Gaussian (sigma_z 0.8-1.5 cm) or uniform (3-6 cm long) profile in z
    centred in [2.5, 5.5] cm, currents 0.1-10 mA (log-uniform).  Keywords
    in `over` replace drawn parameters.  Returns (ParticleGroup, params)."""
    p = dict(kind=kind or ["gaussian", "kv", "waterbag", "mix"][int(rng.integers(4))],
             a=rng.uniform(0.002, 0.006), b=rng.uniform(0.002, 0.006),
             x0=rng.uniform(-0.003, 0.003), y0=rng.uniform(-0.003, 0.003),
             tx=rng.uniform(-0.02, 0.02), ty=rng.uniform(-0.02, 0.02),
             zprof="gauss" if rng.random() < 0.5 else "flat",
             zc=rng.uniform(0.025, 0.055), sz=rng.uniform(0.008, 0.015),
             zlen=rng.uniform(0.03, 0.06), current=10 ** rng.uniform(-4, -2))
    p.update(over)

    def transverse(kind, n):
        if kind == "gaussian":
            return rng.normal(0., 1., n), rng.normal(0., 1., n)
        if kind == "kv":
            r = np.sqrt(rng.random(n))
            th = 2 * np.pi * rng.random(n)
            return r * np.cos(th), r * np.sin(th)
        # waterbag: uniform in the 4-D unit ball, position part
        v = rng.normal(0., 1., (n, 4))
        v *= (rng.random(n) ** 0.25 / np.linalg.norm(v, axis=1))[:, None]
        return v[:, 0], v[:, 1]

    if p["kind"] == "mix":
        k1, k2 = rng.choice(["gaussian", "kv", "waterbag"], 2, replace=False)
        n1 = npart // 2
        u1, v1 = transverse(k1, n1)
        u2, v2 = transverse(k2, npart - n1)
        u, v = np.concatenate([u1, u2]), np.concatenate([v1, v2])
        p["mix"] = "%s+%s" % (k1, k2)
    else:
        u, v = transverse(p["kind"], npart)
    if p["zprof"] == "gauss":
        z = p["zc"] + rng.normal(0., p["sz"], npart)
        zlen = math.sqrt(2 * math.pi) * p["sz"]
    else:
        z = p["zc"] + p["zlen"] * (rng.random(npart) - 0.5)
        zlen = p["zlen"]
    x = p["x0"] + p["a"] * u + p["tx"] * (z - p["zc"])
    y = p["y0"] + p["b"] * v + p["ty"] * (z - p["zc"])
    ux = rng.normal(0., vspread * vz, npart) + p["tx"] * vz
    uy = rng.normal(0., vspread * vz, npart) + p["ty"] * vz
    uz = vz + rng.normal(0., vspread * vz, npart)
    x, y, z, ux, uy, uz = _cull(x, y, z, ux, uy, uz, bounds)
    pg = ParticleGroup(device)
    js = pg.addspecies(charge, mass, p["current"] * zlen / (vz * charge * npart))
    pg.addparticles(js, x, y, z, ux, uy, uz)
    return pg, p
