"""The benchmark problems: mesh, beam and conductors of each configuration.

Common to all: x, y in [-2, 2] cm, z in [0, 8] cm, phi = 0 on every mesh
boundary; singly charged ions of 39 amu moving at vz = 1e6 m/s with a
velocity spread of 1e-3 vz per component; seeded loads (seed 20261202 plus
the configuration index); dt = 1e-9 s.

Choices the configuration text leaves open, fixed here:
  * the pipe (ZCylinderOut, r = 1.8 cm, 0 V) is 10 cm long centred at
    z = 4 cm, so it covers the mesh and its ends make no intercepts;
  * the annulus electrode (ZAnnulus, r_in = 1.2 cm, +10 kV, z in [3, 3.5]
    cm) has r_out = 1.8 cm, touching the pipe;
  * the KV-like beam of the 256^2 x 512 problem is uniform in z over
    [0.5, 6.5] cm (its head reaches z = 8 cm after 15 steps of 1 mm) and
    carries 1 mA;
  * the quadrupole channel's rods are 1.5 cm long, one focusing and one
    defocusing rod set per 4 cm period; the 8 cm mesh holds two periods and
    is periodic in z, for the field and the particles (the reference's
    default particle boundary in z), so a long run circulates the beam
    through the channel; the other problems absorb particles on every face.
"""
import math

from .globals import amu, echarge

DOMAIN = (-0.02, 0.02, -0.02, 0.02, 0.0, 0.08)
VZ = 1.0e6
VSPREAD = 1.0e-3
DT = 1.0e-9
MASS = 39 * amu
SEED = 20261202

CONFIGS = {
    "c0": dict(index=0, n=(64, 64, 64), npart=200_000, beam="gaussian", current=1e-3,
               conductors=None),
    "c1": dict(index=1, n=(128, 128, 256), npart=4_000_000, beam="gaussian", current=1e-3,
               conductors="pipe_annulus"),
    "c2": dict(index=2, n=(256, 256, 512), npart=32_000_000, beam="kv", current=1e-3,
               conductors="pipe_annulus"),
    "c4": dict(index=4, n=(256, 256, 512), npart=32_000_000, beam="gaussian", current=5e-3,
               conductors="quads", zperiodic=True),
}


def field_bounds(cfg):
    """Field boundaries (xm, xp, ym, yp, zm, zp): phi = 0 on every face,
    periodic in z for the quadrupole channel."""
    return [0, 0, 0, 0, 2, 2] if cfg.get("zperiodic") else [0] * 6


def conductors(kind, mod):
    """Conductor list of a configuration, built with the classes of module
    mod (warp_amd, or the oracle's restatement for the CPU runs)."""
    if kind is None:
        return []
    if kind == "pipe_annulus":
        return [mod.ZCylinderOut(0.018, 0.10, 0., zcent=0.04, condid=1),
                mod.ZAnnulus(0.012, 0.018, 0.005, 1.0e4, zcent=0.0325, condid=2)]
    if kind == "quads":
        out = [mod.ZCylinderOut(0.018, 0.10, 0., zcent=0.04, condid=1)]
        cid = 2
        for k in range(2):
            z0 = 0.01 + 0.04 * k
            for zc, sign in ((z0, +1.), (z0 + 0.02, -1.)):
                v = sign * 1.5e4
                for (xc, yc, vv) in ((0.015, 0., v), (-0.015, 0., v), (0., 0.015, -v),
                                     (0., -0.015, -v)):
                    out.append(mod.ZCylinder(0.005, 0.015, vv, xcent=xc, ycent=yc, zcent=zc,
                                             condid=cid))
                    cid += 1
        return out
    raise ValueError(kind)


def set_particle_boundaries(cfg):
    """Particle boundaries of a configuration: absorbing on every face,
    except for the quadrupole channel, whose 8 cm mesh holds two lattice
    periods and is periodic in z (fields and particles)."""
    from .globals import top
    top.pboundxy = 0
    top.pbound0 = top.pboundnz = 2 if cfg.get("zperiodic") else 0


def load_beam(cfg, device="cuda", npart=None):
    """Seeded particle load of a configuration (returns ParticleGroup)."""
    from .particles import load_gaussian, load_kv_ellipse
    npart = cfg["npart"] if npart is None else npart
    seed = SEED + cfg["index"]
    if cfg["beam"] == "gaussian":
        pg, js, n = load_gaussian(npart, (0.004, 0.004, 0.015), 0.04, VZ, VSPREAD, DOMAIN, seed,
                                  charge=echarge, mass=MASS, current=cfg["current"],
                                  device=device)
    else:
        pg, js, n = load_kv_ellipse(npart, 0.006, 0.006, 0.005, 0.065, VZ, VSPREAD, DOMAIN, seed,
                                    charge=echarge, mass=MASS, current=cfg["current"],
                                    device=device)
    return pg


def make_solver(cfg, precision="fp64", device="cuda", **kw):
    """The registered-solver set-up of a configuration."""
    import warp_amd
    from .multigrid import MultiGrid3D
    nx, ny, nz = cfg["n"]
    x0, x1, y0, y1, z0, z1 = DOMAIN
    s = MultiGrid3D(nx=nx, ny=ny, nz=nz, xmmin=x0, xmmax=x1, ymmin=y0, ymmax=y1, zmmin=z0,
                    zmmax=z1, bounds=field_bounds(cfg), precision=precision, device=device,
                    **kw)
    for c in conductors(cfg["conductors"], warp_amd):
        s.installconductor(c)
    return s
