"""Time-step driver restating the reference's electrostatic step3d
(pywarp90/w3d.F:676-1029) for registered solvers: padvnc3d "fullv"
(fetch E, two half kicks, drift), particle boundaries, charge deposition
(loadrho) and the field solve, then time bookkeeping.  ``generate`` does
the initial deposit and solve (w3dgen -> step3d("w3dgen")).
"""
import time

import torch

from . import fieldsolver as FS
from .globals import top, w3d


def generate():
    """Initial loadrho and field solve (scripts/ctl.py:65-75)."""
    FS.loadrhoregistered()
    FS.fieldsolregistered()


def particleboundaries3d(pg, solver, nlost=None):
    """Absorbing particle boundaries (particleboundaries3d, w3d.F:11292-):
    particles leaving the mesh are removed.  nlost, when known (counted by
    the fused advance), skips the scan if no particle left."""
    if nlost == 0:
        return
    x, y, z = pg.xp, pg.yp, pg.zp
    keep = ((x >= solver.xmmin) & (x < solver.xmmax) & (y >= solver.ymmin) &
            (y < solver.ymmax) & (z >= solver.zmmin + top.zgrid) &
            (z < solver.zmmax + top.zgrid))
    if not bool(keep.all()):
        pg.compact(keep)


def padvnc3d_fullv(pg, dt):
    """padvnc3d("fullv") for B = 0 (w3d.F:1780-1986)."""
    solvers = FS.getregisteredsolvers()
    if len(solvers) == 1 and hasattr(solvers[0], "gather_push"):
        for js in range(pg.ns):
            solvers[0].gather_push(pg, js, dt)
        return
    halfdt = 0.5 * (dt * 1 * 1.)
    from . import _native as N
    for js in range(pg.ns):
        sl = pg.species_slice(js)
        n = sl.stop - sl.start
        if n == 0:
            continue
        for c in ("ex", "ey", "ez"):
            getattr(pg, c)[sl].zero_()
        w3d.pgroupfsapi, w3d.ipminfsapi, w3d.npfsapi, w3d.jsfsapi = pg, sl.start + 1, n, js
        w3d.xfsapi = None
        FS.fetcheregistered()
        w3d.npfsapi = 0
        dev = pg.xp.device
        s = N.stream_of(dev)
        col = lambda c: getattr(pg, c)[sl].data_ptr()
        for _ in range(2):
            N.call("wm_epush3d", n, col("uxp"), col("uyp"), col("uzp"), col("ex"), col("ey"),
                   col("ez"), pg.sq[js], pg.sm[js], halfdt, s)
        N.call("wm_xpush3d", n, col("xp"), col("yp"), col("zp"), col("uxp"), col("uyp"),
               col("uzp"), col("gaminv"), dt * 1 * 1., s)


def step(n=1, pgroup=None, boundaries=True):
    """Advance n time steps (scripts/ctl.py:78-104 -> w3dexe)."""
    pg = pgroup if pgroup is not None else top.pgroup
    solvers = FS.getregisteredsolvers()
    for _ in range(n):
        t0 = time.perf_counter()
        padvnc3d_fullv(pg, top.dt)
        if boundaries and solvers:
            nlost = None
            if len(solvers) == 1 and hasattr(solvers[0], "take_lost_count"):
                nlost = solvers[0].take_lost_count()
            particleboundaries3d(pg, solvers[0], nlost)
        if (top.particlesortinterval > 0 and solvers and hasattr(solvers[0], "sortparticles")
                and top.it % top.particlesortinterval == 0):
            solvers[0].sortparticles(pg)
        FS.loadrhoregistered()
        FS.fieldsolregistered()
        top.it += 1
        top.time += top.dt
        top.steptime += time.perf_counter() - t0
