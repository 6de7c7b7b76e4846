"""The time-stepped quadrupole-channel run (the 1000-step configuration,
here at a mesh and step count the FP64 oracle follows in seconds): the
seeded 5 mA Gaussian beam of the 64^3 recipe circulates through two
periods of the electrostatic quadrupole channel (four rods per cell at
+-15 kV, 4 cm period, rod radius 5 mm at 1.5 cm, in the grounded pipe;
fields and particles periodic in z).  The step driver runs it on the GPU
in FP64 and with the mixed-precision solve plus the extrapolated initial
guess; the same steps run with the FP64 oracle (deposit, solve, gather,
halfv/fullv/synchv pushes, particle boundaries and clearpart).  The RMS
emittances of the special steps are compared: their drift over the run
(eps(N)/eps(0) - 1) is stated for each, and the GPU runs must follow the
oracle's history within the bounds asserted below."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX, NY, NZ = 32, 32, 64
NSTEPS = 160          # two passes of the beam through the 8 cm channel
NHIST = 10
NPART = 20000


def _oracle_history(cfg, x, y, z, u, q, m, w, dt):
    """FP64 oracle run of the step driver's sequence (generate, then
    NSTEPS steps with the special-step cadence of nhist and the last step),
    z-periodic; returns {it: (epsx, epsy)} at the history steps."""
    from warp_amd import configs, diagnostics
    x0, x1, y0, y1, z0, z1 = configs.DOMAIN
    dx, dy, dz = (x1 - x0) / NX, (y1 - y0) / NY, (z1 - z0) / NZ
    bounds = configs.field_bounds(cfg)
    co = O.install_conductors(configs.conductors(cfg["conductors"], O), NX, NY, NZ, x0, x1,
                              y0, y1, z0, z1)
    phi = np.zeros((NZ + 3, NY + 3, NX + 3))
    gi = np.ones(len(x))

    def solve():
        rho = np.zeros((NZ + 1, NY + 1, NX + 1))
        O.setrho3d(rho, x, y, z, q, w, dx, dy, dz, x0, y0, z0)
        O.applyrhoboundaryconditions3d(rho, bounds)
        O.multigrid3dsolve(phi, rho, dx, dy, dz, bounds=bounds, conductors=co)

    def efield():
        e = [np.zeros(len(x)) for _ in range(3)]
        O.sete3d(phi, x, y, z, *e, dx, dy, dz, x0, y0, z0)
        return e

    def moments():
        mm = diagnostics.finalize(list(O.getzmmnt(x, y, z, *u, np.ones(len(x)))))
        return mm["epsx"], mm["epsy"]

    solve()
    hist = {0: moments()}
    special = True
    for it in range(1, NSTEPS + 1):
        e = efield()
        for _ in range(1 if special else 2):
            O.epush3d(*u, *e, q, m, 0.5 * (dt * 1 * 1.))
        O.xpush3d(x, y, z, *u, gi, dt * 1 * 1.)
        O.particleboundaries3d(x, y, z, gi, dx, dy, dz, NX, NY, NZ, x0, y0, z0, 0.0,
                               zperiodic=True)
        n = O.clearpart([x, y, z] + list(u) + [gi], 1)
        x, y, z, gi = x[:n].copy(), y[:n].copy(), z[:n].copy(), gi[:n].copy()
        u = [a[:n].copy() for a in u]
        solve()
        special = it % NHIST == 0 or it == NSTEPS
        if special:
            e = efield()
            O.epush3d(*u, *e, q, m, 0.5 * (dt * 1 * 1.))
            if it % NHIST == 0:
                hist[it] = moments()
    return hist, len(x)


def _gpu_history(device, cfg, precision, guess):
    import warp_amd
    from warp_amd import configs, stepper
    from warp_amd.globals import top
    pg = configs.load_beam(cfg, device, npart=NPART)
    c = dict(cfg, n=(NX, NY, NZ))
    s = configs.make_solver(c, precision, device, mgverbose=-1, guess=guess)
    saved = {k: getattr(top, k) for k in ("pgroup", "dt", "it", "time", "nhist", "lspecial",
                                          "particlesortinterval", "pbound0", "pboundnz",
                                          "pboundxy", "hist")}
    try:
        top.pgroup, top.dt, top.it, top.time, top.nhist = pg, configs.DT, 0, 0., NHIST
        top.particlesortinterval, top.hist = 0, []
        configs.set_particle_boundaries(cfg)
        warp_amd.registersolver(s)
        stepper.generate()
        h0 = (top.moments["epsx"], top.moments["epsy"])
        stepper.step(NSTEPS)
        hist = {0: h0}
        for rec in top.hist:
            hist[rec["it"]] = (rec["epsx"], rec["epsy"])
        return hist, pg.nps[0], list(s.mgiters_history)
    finally:
        warp_amd.unregistersolver(s)
        for k, v in saved.items():
            setattr(top, k, v)


def test_quadrupole_channel_emittance_drift(device):
    from warp_amd import configs
    cfg = configs.CONFIGS["c4"]
    pg = configs.load_beam(cfg, "cpu", npart=NPART)
    cols = [getattr(pg, c).numpy().copy() for c in ("xp", "yp", "zp", "uxp", "uyp", "uzp")]
    ho, no = _oracle_history(cfg, *cols[:3], cols[3:], pg.sq[0], pg.sm[0], pg.sw[0],
                             configs.DT)
    h64, n64, it64 = _gpu_history(device, cfg, "fp64", "previous")
    hmx, nmx, itmx = _gpu_history(device, cfg, "bf16", "linear")
    steps = sorted(ho)
    assert steps == sorted(h64) == sorted(hmx) and steps[-1] == NSTEPS
    drift = lambda h, k: h[NSTEPS][k] / h[0][k] - 1.
    print("\nquadrupole channel %dx%dx%d, %d particles (%d/%d/%d at the end), %d steps: "
          "emittance drift epsx/epsy oracle %+.6e/%+.6e, GPU FP64 %+.6e/%+.6e, "
          "GPU bf16+linear %+.6e/%+.6e; V-cycles per solve FP64 %.2f, bf16+linear %.2f"
          % (NX, NY, NZ, NPART, no, n64, nmx, NSTEPS, drift(ho, 0), drift(ho, 1),
             drift(h64, 0), drift(h64, 1), drift(hmx, 0), drift(hmx, 1),
             np.mean(it64[1:]), np.mean(itmx[1:])))
    assert no == n64 == nmx
    for k in (0, 1):
        # the emittance changes over the run (the channel focuses)
        assert abs(drift(ho, k)) > 1e-4
        for it in steps:
            # FP64: only the deposit's order of summation differs
            assert ho[it][k] == pytest.approx(h64[it][k], rel=1e-9), (it, k)
            # mixed precision: every solve still reaches mgtol = 1e-6 V
            assert ho[it][k] == pytest.approx(hmx[it][k], rel=1e-6), (it, k)
        assert abs(drift(h64, k) - drift(ho, k)) <= 1e-9
        assert abs(drift(hmx, k) - drift(ho, k)) <= 1e-6
