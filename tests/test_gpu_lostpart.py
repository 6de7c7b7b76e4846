"""GPU parity of the absorbing particle boundaries and lost-particle
removal (particleboundaries3d -> processlostpart -> clearpart, fill
methods 1 and 2) and of getgrid3d (MultiGrid3D.fetchphi) against the
FP64 oracle, bit for bit, called through the C ABI."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX, NY, NZ = 16, 12, 20
LO = (-0.02, -0.015, 0.0)
DX, DY, DZ = 0.04 / NX, 0.03 / NY, 0.08 / NZ
NAMES = ("xp", "yp", "zp", "uxp", "uyp", "uzp", "gaminv", "ex", "ey", "ez")


def _particles(n, seed, lost_frac, where="random"):
    """Columns of n particles; about lost_frac of them outside the mesh (or
    exactly on a face), placed at random, at the front or at the back."""
    rng = np.random.default_rng(seed)
    hi = (LO[0] + NX * DX, LO[1] + NY * DY, LO[2] + NZ * DZ)
    x = rng.uniform(LO[0], hi[0], n)
    y = rng.uniform(LO[1], hi[1], n)
    z = rng.uniform(LO[2], hi[2], n)
    lost = rng.random(n) < lost_frac
    if where == "front":
        lost = np.arange(n) < int(lost_frac * n)
    elif where == "back":
        lost = np.arange(n) >= n - int(lost_frac * n)
    idx = np.nonzero(lost)[0]
    kind = rng.integers(0, 8, len(idx))
    # out in x/y/z, or exactly on a lower/upper face (the reference absorbs
    # x <= xmin and x >= xmax)
    x[idx[kind == 0]] = LO[0] - 1e-4
    x[idx[kind == 1]] = hi[0]
    y[idx[kind == 2]] = LO[1]
    y[idx[kind == 3]] = hi[1] + 1e-3
    z[idx[kind == 4]] = LO[2]
    z[idx[kind == 5]] = hi[2]
    z[idx[kind == 6]] = hi[2] + 0.01
    x[idx[kind == 7]] = LO[0]
    cols = [x, y, z] + [rng.standard_normal(n) for _ in range(3)] + [np.ones(n)] + \
        [rng.standard_normal(n) for _ in range(3)]
    return cols


def _gpu_cols(cols, device):
    return [torch.from_numpy(c.copy()).to(device) for c in cols]


def _ptrs(ts):
    return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


@pytest.mark.parametrize("n,frac,where", [
    (1, 0.0, "random"), (1, 1.0, "random"), (1023, 0.3, "random"), (1024, 0.5, "random"),
    (1025, 0.01, "random"), (5000, 1.0, "random"), (70001, 0.0, "random"),
    (70001, 0.2, "random"), (70001, 0.05, "front"), (70001, 0.05, "back"),
    (262144, 0.6, "random"), (300007, 0.003, "random")])
@pytest.mark.parametrize("fillmethod", [1, 2])
def test_particleboundaries_and_clearpart_match_oracle(device, n, frac, where, fillmethod):
    from warp_amd import _native as N
    cols = _particles(n, 7 + n, frac, where)
    g = _gpu_cols(cols, device)
    work = torch.empty(int(N.lib().wm_clearpart_worksize(n)), dtype=torch.uint8, device=device)
    out = torch.full((1,), -1, dtype=torch.int64, device=device)
    dst = [torch.full_like(t, np.nan) for t in g] if fillmethod == 2 else None
    N.call("wm_particleboundaries3d", n, len(g), _ptrs(g), LO[0], LO[1], LO[2], DX, DY, DZ,
           NX, NY, NZ, 0.0, 0, fillmethod, _ptrs(dst) if dst else None, out.data_ptr(),
           work.data_ptr(), N.stream_of(device))
    live = int(out.item())
    O.particleboundaries3d(*cols[:3], cols[6], DX, DY, DZ, NX, NY, NZ, *LO)
    want = O.clearpart(cols, fillmethod)
    assert live == want
    res = dst if fillmethod == 2 else g
    for name, a, b in zip(NAMES, res, cols):
        np.testing.assert_array_equal(a[:live].cpu().numpy(), b[:want], err_msg=name)
    if fillmethod == 1:
        # the moved-from slots are cleared as clearpart clears them
        for k in (5, 6):
            np.testing.assert_array_equal(g[k].cpu().numpy(), cols[k], err_msg=NAMES[k])


@pytest.mark.parametrize("zg", [0.0, 0.013])
def test_periodic_z_wraps_and_absorbs_in_x_y(device, zg):
    """zbound = 2 (pbound0 = pboundnz = periodic): z wraps by whole periods
    (particles several periods out included), x and y still absorb; bitwise
    equal to the oracle."""
    from warp_amd import _native as N
    n = 40000
    cols = _particles(n, 21, 0.2)
    rng = np.random.default_rng(22)
    zlen = NZ * DZ
    k = rng.random(n)
    cols[2][k < 0.1] -= zlen * rng.integers(1, 4, int((k < 0.1).sum()))
    cols[2][k > 0.9] += zlen * rng.integers(0, 3, int((k > 0.9).sum()))
    cols[2] += zg
    g = _gpu_cols(cols, device)
    work = torch.empty(int(N.lib().wm_clearpart_worksize(n)), dtype=torch.uint8, device=device)
    out = torch.zeros(1, dtype=torch.int64, device=device)
    N.call("wm_particleboundaries3d", n, len(g), _ptrs(g), LO[0], LO[1], LO[2], DX, DY, DZ,
           NX, NY, NZ, zg, 2, 1, None, out.data_ptr(), work.data_ptr(), N.stream_of(device))
    O.particleboundaries3d(*cols[:3], cols[6], DX, DY, DZ, NX, NY, NZ, *LO, zg, zperiodic=True)
    want = O.clearpart(cols, 1)
    assert int(out.item()) == want
    for name, a, b in zip(NAMES, g, cols):
        np.testing.assert_array_equal(a[:want].cpu().numpy(), b[:want], err_msg=name)
    z = cols[2][:want] - zg
    assert z.min() >= LO[2] - 1e-12 and z.max() < LO[2] + zlen


def test_clearpart_of_flagged_particles(device):
    """wm_clearpart alone removes what gaminv <= 0 flags (some flags
    negative, as lcountaslost false would leave them)."""
    from warp_amd import _native as N
    rng = np.random.default_rng(5)
    n = 100000
    cols = _particles(n, 5, 0.0)
    flag = rng.random(n)
    cols[6][flag < 0.1] = 0.
    cols[6][flag > 0.95] = -1.
    for method in (1, 2):
        c = [a.copy() for a in cols]
        g = _gpu_cols(c, device)
        work = torch.empty(int(N.lib().wm_clearpart_worksize(n)), dtype=torch.uint8,
                           device=device)
        out = torch.zeros(1, dtype=torch.int64, device=device)
        dst = [torch.empty_like(t) for t in g] if method == 2 else None
        N.call("wm_clearpart", n, len(g), _ptrs(g), method, _ptrs(dst) if dst else None,
               out.data_ptr(), work.data_ptr(), N.stream_of(device))
        want = O.clearpart(c, method)
        live = int(out.item())
        assert live == want
        for a, b in zip(dst if dst else g, c):
            np.testing.assert_array_equal(a[:live].cpu().numpy(), b[:live])


def test_step_removes_lost_particles_on_the_gpu(device):
    """ParticleGroup + the step driver's boundary pass: the fused advance
    counts the particles it pushed out, the boundary pass removes exactly
    those with the reference's fill method, and the species count follows."""
    import warp_amd
    from warp_amd import stepper
    from warp_amd.globals import top
    from warp_amd.particles import ParticleGroup
    cols = _particles(50000, 11, 0.0)
    pg = ParticleGroup(device)
    js = pg.addspecies(1.602176565e-19, 39 * 1.660538921e-27, 1.0)
    pg.addparticles(js, *[torch.from_numpy(c) for c in cols[:6]])
    # a drift that takes particles through every face
    s = warp_amd.MultiGrid3D(nx=NX, ny=NY, nz=NZ, xmmin=LO[0], xmmax=LO[0] + NX * DX,
                             ymmin=LO[1], ymmax=LO[1] + NY * DY, zmmin=LO[2],
                             zmmax=LO[2] + NZ * DZ, bounds=[0] * 6, mgverbose=-1)
    pg.uxp.mul_(2e5)
    pg.uyp.mul_(2e5)
    pg.uzp.mul_(5e5)
    x0 = [getattr(pg, c).cpu().numpy().copy() for c in NAMES]
    s.gather_push(pg, 0, 1.0e-8, "fullv")
    nlost = s.take_lost_count()
    after = [getattr(pg, c).cpu().numpy().copy() for c in NAMES]
    saved = (top.pbound0, top.pboundnz)
    top.pbound0 = top.pboundnz = 0
    try:
        stepper.particleboundaries3d(pg, s, nlost)
    finally:
        top.pbound0, top.pboundnz = saved
    assert nlost > 1000
    assert pg.nps[0] == 50000 - nlost
    O.particleboundaries3d(*after[:3], after[6], s.dx, s.dy, s.dz, NX, NY, NZ, s.xmmin,
                           s.ymmin, s.zmmin)
    n = O.clearpart(after, top.clearlostpart)
    assert n == pg.nps[0]
    for name, a in zip(NAMES, after):
        np.testing.assert_array_equal(getattr(pg, name)[:n].cpu().numpy(), a[:n], err_msg=name)
    assert x0[0].shape == (50000,)


@pytest.mark.parametrize("l2,l4", [(False, False), (True, False), (False, True)])
def test_fetchphi_matches_getgrid3d(device, l2, l4):
    """MultiGrid3D.fetchphi (getgrid3d over phi with guard cells,
    multigrid.py:560-575): bitwise equal to the oracle, including points in
    the guard cells, on the upper faces, outside the guarded mesh (left
    unchanged) and on nodes."""
    import warp_amd
    from warp_amd.globals import w3d
    s = warp_amd.MultiGrid3D(nx=NX, ny=NY, nz=NZ, xmmin=LO[0], xmmax=LO[0] + NX * DX,
                             ymmin=LO[1], ymmax=LO[1] + NY * DY, zmmin=LO[2],
                             zmmax=LO[2] + NZ * DZ, bounds=[0] * 6, mgverbose=-1)
    s.l2symtry, s.l4symtry = l2, l4
    rng = np.random.default_rng(9)
    phi = rng.standard_normal((NZ + 3, NY + 3, NX + 3))
    s.phi = torch.from_numpy(phi).to(device)
    n = 20000
    x = rng.uniform(LO[0] - 2 * DX, LO[0] + (NX + 2) * DX, n)
    y = rng.uniform(LO[1] - 2 * DY, LO[1] + (NY + 2) * DY, n)
    z = rng.uniform(LO[2] - 2 * DZ, LO[2] + (NZ + 2) * DZ, n)
    x[:50] = LO[0] + DX * rng.integers(-1, NX + 2, 50)     # on nodes
    z[50:60] = LO[2] + (NZ + 1) * DZ                        # upper guard face
    f0 = rng.standard_normal(n)
    t = [torch.from_numpy(a.copy()).to(device) for a in (x, y, z, f0)]
    w3d.xfsapi, w3d.yfsapi, w3d.zfsapi, w3d.phifsapi, w3d.npfsapi = t[0], t[1], t[2], t[3], n
    try:
        s.fetchphi()
    finally:
        w3d.xfsapi = w3d.yfsapi = w3d.zfsapi = w3d.phifsapi = None
        w3d.npfsapi = 0
    want = f0.copy()
    O.getgrid3d(x, y, z, want, phi, LO[0] - DX, LO[0] + NX * DX + DX, LO[1] - DY,
                LO[1] + NY * DY + DY, LO[2] - DZ, LO[2] + NZ * DZ + DZ, l2, l4)
    np.testing.assert_array_equal(t[3].cpu().numpy(), want)
    assert (want != f0).sum() > n // 3 and (want == f0).sum() > 0
