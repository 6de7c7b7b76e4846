"""The step driver end to end (generate, then step(7) with the reference's
special-step cadence) on the GPU against the same sequence run with the
FP64 oracle: deposit, multigrid solve, E gather, halfv/fullv/synchv pushes,
absorbing boundaries and the beam moments; in free space, and with the
pipe and annulus installed and the conductor-corrected gather
(f3d.lcorrectede).  The deposit sums in another
order on the GPU (atomics), so the states agree to round-off, not bit for
bit: positions and velocities to 1e-10 relative, the moments to 1e-9."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

NX, NY, NZ = 32, 32, 64
DOM = (-0.02, 0.02, -0.02, 0.02, 0.0, 0.08)


def _oracle_run(x, y, z, u, q, m, w, dt, nsteps, nhist, efetch=1, lcorrectede=False):
    dx, dy, dz = 0.04 / NX, 0.04 / NY, 0.08 / NZ
    lo = (DOM[0], DOM[2], DOM[4])
    phi = np.zeros((NZ + 3, NY + 3, NX + 3))
    co = cf = None
    if lcorrectede:
        co = O.install_conductors(H.pipe_and_annulus(O), NX, NY, NZ, *DOM)
        O.checkconductors(co, NX, NY, NZ, dx, dy, dz)
        cf = O.CondField(co.lists(), NX, NY, NZ)

    def solve(x, y, z):
        rho = np.zeros((NZ + 1, NY + 1, NX + 1))
        O.setrho3d(rho, x, y, z, q, w, dx, dy, dz, *lo)
        O.multigrid3dsolve(phi, rho, dx, dy, dz, conductors=co)

    def efield(x, y, z):
        e = [np.zeros(len(x)) for _ in range(3)]
        if lcorrectede:
            cf.efield(phi, dx, dy, dz)
            se = None
            if efetch == 3:
                se = O.getselfe3d(phi, dx, dy, dz)
                cf.fixefield(se)
            O.sete3d_conductor(cf, phi, se, x, y, z, *e, dx, dy, dz, *lo, efetch=efetch)
        elif efetch == 3:
            O.sete_from_e3d(O.getselfe3d(phi, dx, dy, dz), x, y, z, *e, dx, dy, dz, *lo)
        else:
            O.sete3d(phi, x, y, z, *e, dx, dy, dz, *lo)
        return e

    solve(x, y, z)
    special, centers = True, []
    for it in range(1, nsteps + 1):
        e = efield(x, y, z)
        center = "halfv" if special else "fullv"
        centers.append(center)
        for _ in range(1 if special else 2):
            O.epush3d(*u, *e, q, m, 0.5 * (dt * 1 * 1.))
        gi = np.ones(len(x))
        O.xpush3d(x, y, z, *u, gi, dt * 1 * 1.)
        # particleboundaries3d: flag, then clearpart with fillmethod 1
        O.particleboundaries3d(x, y, z, gi, dx, dy, dz, NX, NY, NZ, *lo)
        cols = [x, y, z] + list(u) + [gi]
        n = O.clearpart(cols, 1)
        x, y, z, u = x[:n].copy(), y[:n].copy(), z[:n].copy(), [a[:n].copy() for a in u]
        solve(x, y, z)
        special = it % nhist == 0 or it == nsteps
        if special:
            centers.append("synchv")
            e = efield(x, y, z)
            O.epush3d(*u, *e, q, m, 0.5 * (dt * 1 * 1.))
    return x, y, z, u, centers


@pytest.mark.parametrize("efetch,lcorrectede", [(1, False), (3, False), (1, True), (3, True)])
def test_step_driver_matches_oracle(device, efetch, lcorrectede):
    import warp_amd
    from warp_amd import diagnostics, stepper
    from warp_amd.globals import f3d, top
    from warp_amd.particles import load_gaussian
    pg, js, n = load_gaussian(20000, (0.004, 0.004, 0.015), 0.04, 1.0e6, 1.0e-3, DOM, 31,
                              device=device)
    x0 = [getattr(pg, c).cpu().numpy().copy() for c in ("xp", "yp", "zp")]
    u0 = [getattr(pg, c).cpu().numpy().copy() for c in ("uxp", "uyp", "uzp")]
    s = warp_amd.MultiGrid3D(nx=NX, ny=NY, nz=NZ, xmmin=DOM[0], xmmax=DOM[1], ymmin=DOM[2],
                             ymmax=DOM[3], zmmin=DOM[4], zmmax=DOM[5], bounds=[0] * 6,
                             mgverbose=-1)
    if lcorrectede:
        for c in H.pipe_and_annulus(warp_amd):
            s.installconductor(c)
    saved_lc = f3d.lcorrectede
    saved = dict(pgroup=top.pgroup, dt=top.dt, it=top.it, time=top.time, nhist=top.nhist,
                 lspecial=top.lspecial, particlesortinterval=top.particlesortinterval,
                 efetch=top.efetch, nzzarr=top.nzzarr, zzmin=top.zzmin, zzmax=top.zzmax,
                 dzz=top.dzz, linechg=top.linechg, ese=top.ese, pbound0=top.pbound0,
                 pboundnz=top.pboundnz)
    calls = []
    orig = s.gather_push

    def spy(pg_, js_, dt_, center="fullv"):
        calls.append(center)
        return orig(pg_, js_, dt_, center)

    s.gather_push = spy
    try:
        top.pgroup, top.dt, top.it, top.time, top.nhist = pg, 1.0e-9, 0, 0., 5
        top.particlesortinterval = 0
        top.pbound0 = top.pboundnz = 0     # absorbing in z, as the oracle run
        top.efetch = efetch
        f3d.lcorrectede = lcorrectede
        warp_amd.registersolver(s)
        stepper.generate()
        stepper.step(7)
        xo, yo, zo, uo, centers = _oracle_run(*[a.copy() for a in x0], [a.copy() for a in u0],
                                              pg.sq[0], pg.sm[0], pg.sw[0], 1.0e-9, 7, 5,
                                              efetch, lcorrectede)
        assert calls == centers
        assert centers == ["halfv", "fullv", "fullv", "fullv", "fullv", "synchv", "halfv",
                           "fullv", "synchv"]
        nlive = pg.nps[0]
        got = [getattr(pg, c)[:nlive].cpu().numpy() for c in ("xp", "yp", "zp", "uxp", "uyp",
                                                               "uzp")]
        assert len(got[0]) == len(xo)
        for a, b in zip(got, [xo, yo, zo] + uo):
            scale = np.abs(b).max()
            assert np.abs(a - b).max() <= 1e-10 * scale
        mm = diagnostics.finalize(list(O.getzmmnt(xo, yo, zo, *uo, np.ones(len(xo)))))
        for key in ("xrms", "yrms", "zrms", "epsx", "epsy", "vzbar"):
            assert top.moments[key] == pytest.approx(mm[key], rel=1e-9), key
        assert top.it == 7 and top.lspecial
        # field diagnostics of the last (special) step, on the solver's rho
        # and phi (w3d.F:997-1003)
        rho, phi = s.rho.cpu().numpy(), s.phi.cpu().numpy()
        lco, _ = O.gtlchg3d(rho, s.dx, s.dy, s.dz, top.zgrid, s.zmmin, top.zzmin, top.zbeam,
                            top.dzz, top.nzzarr)
        assert np.array_equal(top.linechg.cpu().numpy(), lco)
        assert top.ese == O.getese3d(rho, phi, s.dx, s.dy, s.dz)
        assert top.hist[-1]["it"] == 5 and top.hist[-1]["ese"] > 0.
    finally:
        f3d.lcorrectede = saved_lc
        warp_amd.unregistersolver(s)
        for k, v in saved.items():
            setattr(top, k, v)


def test_phi_extrapolate_kernel(device):
    """wm_phi_extrapolate: phi <- 2 phi - prev, prev <- old phi, bitwise."""
    from warp_amd import _native as N
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(100003), rng.standard_normal(100003)
    pa, pb = torch.from_numpy(a.copy()).to(device), torch.from_numpy(b.copy()).to(device)
    N.call("wm_phi_extrapolate", pa.data_ptr(), pb.data_ptr(), pa.numel(), N.stream_of(device))
    assert np.array_equal(pa.cpu().numpy(), 2. * a - b)
    assert np.array_equal(pb.cpu().numpy(), a)


def test_linear_guess_reaches_mgtol(device):
    """The step driver with the initial guess extrapolated from the two
    previous solves: every solve still stops on mgtol, the particles end
    where the reference's guess (the previous phi) takes them up to the
    solves' tolerance, and the extrapolated guess needs no more V-cycles."""
    import warp_amd
    from warp_amd import stepper
    from warp_amd.globals import top
    from warp_amd.particles import load_gaussian
    out = {}
    for guess in ("previous", "linear"):
        pg, js, n = load_gaussian(20000, (0.004, 0.004, 0.015), 0.04, 1.0e6, 1.0e-3, DOM, 31,
                                  device=device)
        s = warp_amd.MultiGrid3D(nx=NX, ny=NY, nz=NZ, xmmin=DOM[0], xmmax=DOM[1], ymmin=DOM[2],
                                 ymmax=DOM[3], zmmin=DOM[4], zmmax=DOM[5], bounds=[0] * 6,
                                 mgverbose=-1, guess=guess)
        for c in H.pipe_and_annulus(warp_amd):
            s.installconductor(c)
        saved = dict(pgroup=top.pgroup, dt=top.dt, it=top.it, time=top.time,
                     particlesortinterval=top.particlesortinterval, pbound0=top.pbound0,
                     pboundnz=top.pboundnz)
        try:
            top.pgroup, top.dt, top.it, top.time = pg, 1.0e-9, 0, 0.
            top.particlesortinterval = 0
            top.pbound0 = top.pboundnz = 0
            warp_amd.registersolver(s)
            stepper.generate()
            stepper.step(8)
            errs = s.mgerror
            out[guess] = ([getattr(pg, c)[:pg.nps[0]].cpu().numpy() for c in ("xp", "zp", "uzp")],
                          list(s.mgiters_history), errs)
        finally:
            warp_amd.unregistersolver(s)
            for k, v in saved.items():
                setattr(top, k, v)
    (a, ha, ea), (b, hb, eb) = out["previous"], out["linear"]
    print("\nV-cycles per solve: previous phi %s, extrapolated %s" % (ha, hb))
    assert ea <= 1e-6 and eb <= 1e-6
    assert sum(hb[3:]) <= sum(ha[3:])
    for u, v in zip(a, b):
        assert np.abs(u - v).max() <= 1e-9 * np.abs(u).max()
