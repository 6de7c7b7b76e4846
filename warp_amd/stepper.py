"""Time-step driver restating the reference's electrostatic step3d
(pywarp90/w3d.F:511-1029) for registered solvers.  Each step: padvnc3d
"halfv" when the previous step was special (velocities at the integer
time level) else "fullv" (fetch E, half kicks, drift), particle boundaries,
charge deposition (loadrho) and the field solve; a special step then ends
with padvnc3d "synchv" (fetch E at the new positions, half kick) and the
beam moments (getzmmnt) and the field diagnostics (line charge top.linechg,
electrostatic energy top.ese).  ``generate`` does the initial deposit and
solve and the diagnostics of the loaded beam (w3dgen -> step3d("w3dgen"),
"gen").
"""
import time

import torch

from . import diagnostics
from . import fieldsolver as FS
from .globals import top, w3d

MACHEPS = 1.0e-14  # pywarp90/top.h


def thisstep(it, icontrol):
    """dtop.F:764-789: icontrol = (start, end, increment, extra steps...)."""
    istart, iend, iinc = icontrol[0], icontrol[1], icontrol[2]
    if iinc != 0 and istart <= it <= iend and (it - istart) % iinc == 0:
        return True
    return any(it == i for i in icontrol[3:])


def special_step(llast):
    """The lspecial flag of step3d (w3d.F:854-880) for the current top.it,
    top.time: end of run, moments, history, dump or last step of this call
    (the plot controls, always off here, are not modelled)."""
    lfinishd = (top.it >= top.nt) or (top.time >= top.tstop * (1. - MACHEPS)) or \
        (top.zbeam >= top.zstop)
    lmoments = thisstep(top.it, top.itmomnts)
    lhist = top.nhist != 0 and top.it % top.nhist == 0
    ldump = top.it % top.itdump == 0
    return bool(lfinishd or lmoments or lhist or ldump or llast or top.it == 0 or top.allspecl), lhist


def generate():
    """Initial loadrho and field solve (scripts/ctl.py:65-75), then the
    "gen" phase: E fetched at the loaded positions, velocities unchanged,
    moments of the loaded beam; the first step is then a halfv step."""
    FS.loadrhoregistered()
    FS.fieldsolregistered()
    pg = top.pgroup
    if pg is not None and pg.nplive > 0:
        padvnc3d("gen", pg, top.dt)
        top.moments = diagnostics.getzmmnt(pg)
    field_diagnostics(special=True)
    top.lspecial = True


def field_diagnostics(special):
    """The field diagnostics at the end of step3d (w3d.F:987-1002): rhodia
    on every step; line charge, on-axis rho, electrostatic energy, on-axis
    phi and Ez on special steps."""
    if not FS.getregisteredsolvers():
        return
    FS.rhodiaregistered()
    if special:
        FS.gtlchgregistered()
        FS.srhoaxregistered()
        FS.geteseregistered()
        FS.sphiaxregistered()
        FS.sezaxregistered()


def particleboundaries3d(pg, solver, nlost=None):
    """Absorbing particle boundaries (particleboundaries3d,
    w3d.F:11292-11450): particles outside the field mesh in x or y (and in
    z unless top.pbound0 is periodic, when z wraps instead) are flagged lost
    and removed by processlostpart -> clearpart with top.clearlostpart's
    fill method, all on the GPU (the host reads back the new count).
    nlost, when known (the fused advance counts the particles that crossed
    any face of the mesh), skips the pass if none did."""
    if nlost == 0:
        return
    zb = top.pbound0
    if top.pboundnz != zb or zb not in (0, 2) or top.pboundxy != 0:
        raise NotImplementedError(
            "particle boundaries: absorbing x/y (pboundxy = absorb) with absorbing or "
            "periodic z (pbound0 = pboundnz) only; got pbound0=%s pboundnz=%s pboundxy=%s"
            % (top.pbound0, top.pboundnz, top.pboundxy))
    mesh = (solver.xmmin, solver.ymmin, solver.zmmin, solver.dx, solver.dy, solver.dz,
            solver.nx, solver.ny, solver.nz, top.zgrid, zb)
    for js in range(pg.ns):
        pg._remove_lost(js, top.clearlostpart, mesh)


def padvnc3d(center, pg, dt):
    """padvnc3d(center) for B = 0 (w3d.F:1575-2330): "fullv" (two half
    kicks, drift), "halfv" (one half kick, drift), "synchv" (one half
    kick), "gen" (E fetched only)."""
    solvers = FS.getregisteredsolvers()
    if center != "gen" and len(solvers) == 1 and hasattr(solvers[0], "gather_push"):
        for js in range(pg.ns):
            solvers[0].gather_push(pg, js, dt, center)
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
        if center == "gen":
            continue
        dev = pg.xp.device
        s = N.stream_of(dev)
        col = lambda c: getattr(pg, c)[sl].data_ptr()
        for _ in range(2 if center == "fullv" else 1):
            N.call("wm_epush3d", n, col("uxp"), col("uyp"), col("uzp"), col("ex"), col("ey"),
                   col("ez"), pg.sq[js], pg.sm[js], halfdt, s)
        if center != "synchv":
            N.call("wm_xpush3d", n, col("xp"), col("yp"), col("zp"), col("uxp"), col("uyp"),
                   col("uzp"), col("gaminv"), dt * 1 * 1., s)


def padvnc3d_fullv(pg, dt):
    """padvnc3d("fullv")."""
    padvnc3d("fullv", pg, dt)


_beforestep, _afterstep = [], []


def installbeforestep(f):
    """Calls f() before every step (controllers.installbeforestep,
    scripts/controllers.py:543-552)."""
    _beforestep.append(f)


def uninstallbeforestep(f):
    _beforestep.remove(f)


def installafterstep(f):
    """Calls f() after every step (controllers.installafterstep,
    scripts/controllers.py:557-566)."""
    _afterstep.append(f)


def uninstallafterstep(f):
    _afterstep.remove(f)


def step(n=1, pgroup=None, boundaries=True):
    """Advance n time steps (scripts/ctl.py:78-104 -> w3dexe -> step3d),
    with the before/after-step functions around each."""
    pg = pgroup if pgroup is not None else top.pgroup
    solvers = FS.getregisteredsolvers()
    for k in range(n):
        t0 = time.perf_counter()
        for f in list(_beforestep):
            f()
        top.it += 1
        padvnc3d("halfv" if top.lspecial else "fullv", pg, top.dt)
        top.time += top.dt
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
        top.lspecial, lhist = special_step(llast=(k == n - 1))
        if top.lspecial:
            top.nspecial += 1
            padvnc3d("synchv", pg, top.dt)
            top.moments = diagnostics.getzmmnt(pg)
        field_diagnostics(top.lspecial)
        if top.lspecial and lhist:
            top.hist.append(dict(it=top.it, time=top.time, ese=top.ese, **top.moments))
        for f in list(_afterstep):
            f()
        top.steptime += time.perf_counter() - t0
