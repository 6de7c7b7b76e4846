"""GPU parity of the particle kernels (deposit, gather, push, grid E)
against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

LO = (-0.02, -0.02, 0.0)
NX, NY, NZ = 64, 64, 64
DX, DY, DZ = 0.04 / NX, 0.04 / NY, 0.08 / NZ


def _beam(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 0.004, n)
    y = rng.normal(0, 0.004, n)
    z = 0.04 + rng.normal(0, 0.015, n)
    k = (x > LO[0]) & (x < -LO[0]) & (y > LO[1]) & (y < -LO[1]) & (z > 0) & (z < 0.08)
    return x[k], y[k], z[k]


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def _cell_order(x, y, z):
    """Stable order by CIC cell (x fastest), as after particlesortxyzwithcopy."""
    ix = ((x - LO[0]) * (1. / DX)).astype(np.int64)
    iy = ((y - LO[1]) * (1. / DY)).astype(np.int64)
    iz = ((z - LO[2]) * (1. / DZ)).astype(np.int64)
    return np.argsort(ix + (NX + 1) * (iy + (NY + 1) * iz), kind="stable")


@pytest.mark.parametrize("order", ["random", "sorted_call", "cell_ordered"])
def test_deposit_matches_oracle(device, order):
    """setrho3d: the GPU sums in another order, so rho agrees to rounding
    (relative 1e-13 of the peak) and the total charge is conserved.
    Random order (every lane its own run), a sort inside the call, and
    particles already in cell order (long runs reduced across the wave)."""
    from warp_amd import _native as N
    x, y, z = _beam(200000, 1)
    if order == "cell_ordered":
        p = _cell_order(x, y, z)
        x, y, z = x[p], y[p], z[p]
    sorted_ = order == "sorted_call"
    q, w = 1.602176565e-19, 3.7e4
    rho = torch.zeros((NZ + 1, NY + 1, NX + 1), dtype=torch.float64, device=device)
    args = [rho.data_ptr(), len(x), _t(x, device).data_ptr(), _t(y, device).data_ptr(),
            _t(z, device).data_ptr(), 0.0, q, w, NX, NY, NZ, DX, DY, DZ, LO[0], LO[1], LO[2]]
    keep = [_t(x, device), _t(y, device), _t(z, device)]
    args[2:5] = [k.data_ptr() for k in keep]
    st = N.stream_of(device)
    if sorted_:
        work = torch.empty(N.lib().wm_setrho3d_sorted_worksize(len(x), NX, NY, NZ),
                           dtype=torch.uint8, device=device)
        N.call("wm_setrho3d_sorted", *args, work.data_ptr(), st)
    else:
        ns = C.c_int64(-1)
        N.call("wm_setrho3d", *args, C.byref(ns), st)
        assert ns.value == 0
    ro = np.zeros((NZ + 1, NY + 1, NX + 1))
    O.setrho3d(ro, x, y, z, q, w, DX, DY, DZ, *LO)
    r = rho.cpu().numpy()
    assert np.abs(r - ro).max() <= 1e-13 * np.abs(ro).max()
    total = q * w * len(x) / (DX * DY * DZ)
    assert abs(r.sum() - total) <= 1e-12 * total


def test_deposit_skips_out_of_grid(device):
    from warp_amd import _native as N
    x = np.array([0.0, 0.03, -0.025, 0.0])
    y = np.array([0.0, 0.0, 0.0, 0.0])
    z = np.array([0.04, 0.04, 0.04, 0.08])   # 0.08 = zmmax is outside [0, 0.08)
    rho = torch.zeros((NZ + 1, NY + 1, NX + 1), dtype=torch.float64, device=device)
    keep = [_t(v, device) for v in (x, y, z)]
    ns = C.c_int64(-1)
    N.call("wm_setrho3d", rho.data_ptr(), 4, keep[0].data_ptr(), keep[1].data_ptr(),
           keep[2].data_ptr(), 0.0, 1.0, 1.0, NX, NY, NZ, DX, DY, DZ, *LO, C.byref(ns),
           N.stream_of(device))
    assert ns.value == 3
    assert abs(rho.sum().item() - 1.0 / (DX * DY * DZ)) < 1e-6


def test_deposit_runs_with_outside_particles(device):
    """Cell-ordered particles interleaved with out-of-grid ones: runs are
    cut at every skipped particle and the rest still sums exactly."""
    from warp_amd import _native as N
    x, y, z = _beam(20000, 11)
    p = _cell_order(x, y, z)
    x, y, z = x[p], y[p], z[p]
    xo = x.copy()
    xo[::7] = 0.5            # every 7th particle outside the mesh
    keep = [_t(v, device) for v in (xo, y, z)]
    rho = torch.zeros((NZ + 1, NY + 1, NX + 1), dtype=torch.float64, device=device)
    ns = C.c_int64(-1)
    N.call("wm_setrho3d", rho.data_ptr(), len(x), keep[0].data_ptr(), keep[1].data_ptr(),
           keep[2].data_ptr(), 0.0, 1.0, 2.0, NX, NY, NZ, DX, DY, DZ, *LO, C.byref(ns),
           N.stream_of(device))
    inside = np.ones(len(x), bool)
    inside[::7] = False
    assert ns.value == (~inside).sum()
    ro = np.zeros((NZ + 1, NY + 1, NX + 1))
    O.setrho3d(ro, x[inside], y[inside], z[inside], 1.0, 2.0, DX, DY, DZ, *LO)
    assert np.abs(rho.cpu().numpy() - ro).max() <= 1e-13 * np.abs(ro).max()


@pytest.mark.parametrize("mesh", [(8, 8, 16), (64, 64, 128), (256, 256, 512)])
def test_particle_sort_key_widths(device, mesh):
    """The hand-written radix sort over keys of 10, 20 and 26 bits (2, 3
    and 4 passes of 8-bit digits), 1.3M particles with many equal keys
    (positions drawn on a coarse lattice plus particles outside the mesh):
    the permutation of a stable sort by the reference's block index."""
    from warp_amd import _native as N
    nx, ny, nz = mesh
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    dx, dy, dz = [(b - a) / k for a, b, k in zip(lo, hi, mesh)]
    rng = np.random.default_rng(31)
    n = 1_300_000
    x = lo[0] + np.round(rng.uniform(-0.1, 1.1, n) * 40) / 40 * (hi[0] - lo[0])
    y = lo[1] + rng.uniform(0, 1, n) * (hi[1] - lo[1])
    z = lo[2] + rng.uniform(0, 1, n) * (hi[2] - lo[2])
    fx = np.abs(x - lo[0]) / dx
    fy = np.abs(y - lo[1]) / dy
    fz = np.abs(z - lo[2]) / dz
    ix = np.where(fx < nx, fx.astype(np.int64), nx)
    iy = np.where(fy < ny, fy.astype(np.int64), ny)
    iz = np.where(fz < nz, fz.astype(np.int64), nz)
    key = ix + (nx + 1) * (iy + (ny + 1) * iz)
    want = np.argsort(key, kind="stable")
    perm = torch.empty(n, dtype=torch.int32, device=device)
    work = torch.empty(N.lib().wm_particle_sort_worksize(n, nx, ny, nz), dtype=torch.uint8,
                       device=device)
    g = [_t(v, device) for v in (x, y, z)]
    N.call("wm_particle_sort", n, *[a.data_ptr() for a in g], *lo, dx, dy, dz, nx, ny, nz,
           perm.data_ptr(), work.data_ptr(), N.stream_of(device))
    np.testing.assert_array_equal(perm.cpu().numpy(), want)


def test_particle_sort_matches_counting_sort(device):
    """particlesortxyzwithcopy: the permutation equals a stable sort by the
    reference's block index int(|x-xmmin|/dx) + (nx+1)*(iy + (ny+1)*iz),
    and every column is permuted with it."""
    from warp_amd import _native as N
    from warp_amd.particles import ParticleGroup, particlesortxyzwithcopy
    x, y, z = _beam(300000, 12)
    n = len(x)
    rng = np.random.default_rng(13)
    u = [rng.normal(0, 1e3, n) for _ in range(3)]
    key = ((np.abs(x - LO[0]) / DX).astype(np.int64) +
           (NX + 1) * ((np.abs(y - LO[1]) / DY).astype(np.int64) +
                       (NY + 1) * (np.abs(z - LO[2]) / DZ).astype(np.int64)))
    want = np.argsort(key, kind="stable")
    perm = torch.empty(n, dtype=torch.int32, device=device)
    work = torch.empty(N.lib().wm_particle_sort_worksize(n, NX, NY, NZ), dtype=torch.uint8,
                       device=device)
    g = [_t(v, device) for v in (x, y, z)]
    N.call("wm_particle_sort", n, *[a.data_ptr() for a in g], *LO, DX, DY, DZ, NX, NY, NZ,
           perm.data_ptr(), work.data_ptr(), N.stream_of(device))
    np.testing.assert_array_equal(perm.cpu().numpy(), want)
    # two species: each sorted within its own slice
    pg = ParticleGroup(device)
    h = n // 3
    for a, b in ((0, h), (h, n)):
        js = pg.addspecies(1.0, 1.0, 1.0)
        pg.addparticles(js, x[a:b], y[a:b], z[a:b], u[0][a:b], u[1][a:b], u[2][a:b])
    particlesortxyzwithcopy(pg, DX, DY, DZ, *LO, NX, NY, NZ)
    for a, b in ((0, h), (h, n)):
        p = a + np.argsort(key[a:b], kind="stable")
        np.testing.assert_array_equal(pg.xp[a:b].cpu().numpy(), x[p])
        np.testing.assert_array_equal(pg.zp[a:b].cpu().numpy(), z[p])
        np.testing.assert_array_equal(pg.uyp[a:b].cpu().numpy(), u[1][p])


def _phi(seed):
    return np.random.default_rng(seed).standard_normal((NZ + 3, NY + 3, NX + 3))


def test_gather_bitwise(device):
    """sete3d: bit-identical E at every particle."""
    from warp_amd import _native as N
    x, y, z = _beam(100000, 2)
    phi = _phi(3)
    ex, ey, ez = (torch.full((len(x),), 0.5, dtype=torch.float64, device=device) for _ in range(3))
    keep = [_t(v, device) for v in (x, y, z)]
    N.call("wm_sete3d", _t(phi, device).data_ptr(), len(x), keep[0].data_ptr(),
           keep[1].data_ptr(), keep[2].data_ptr(), 0.0, *LO, DX, DY, DZ, NX, NY, NZ,
           ex.data_ptr(), ey.data_ptr(), ez.data_ptr(), N.stream_of(device))
    eo = [np.full(len(x), 0.5) for _ in range(3)]
    O.sete3d(phi, x, y, z, *eo, DX, DY, DZ, *LO)
    for a, b in zip((ex, ey, ez), eo):
        np.testing.assert_array_equal(a.cpu().numpy(), b)


def test_push_bitwise(device):
    from warp_amd import _native as N
    rng = np.random.default_rng(4)
    n = 50000
    u = [rng.normal(0, 1e3, n) for _ in range(3)]
    e = [rng.normal(0, 1e5, n) for _ in range(3)]
    xx = [rng.normal(0, 1e-2, n) for _ in range(3)]
    gi = np.ones(n)
    ug = [_t(a, device) for a in u]
    eg = [_t(a, device) for a in e]
    xg = [_t(a, device) for a in xx]
    st = N.stream_of(device)
    q, m, dt = 1.602176565e-19, 39 * 1.660538921e-27, 1e-9
    N.call("wm_epush3d", n, *[a.data_ptr() for a in ug], *[a.data_ptr() for a in eg], q, m,
           0.5 * dt, st)
    N.call("wm_xpush3d", n, *[a.data_ptr() for a in xg], *[a.data_ptr() for a in ug],
           _t(gi, device).data_ptr(), dt, st)
    uo = [a.copy() for a in u]
    xo = [a.copy() for a in xx]
    O.epush3d(*uo, *e, q, m, 0.5 * dt)
    O.xpush3d(*xo, *uo, gi, dt)
    for a, b in zip(ug + xg, uo + xo):
        np.testing.assert_array_equal(a.cpu().numpy(), b)


def test_gather_push_matches_sequence(device):
    """The fused fullv advance equals E=0, sete3d, epush3d(dt/2) twice,
    xpush3d(dt) of the oracle, bit for bit."""
    from warp_amd import _native as N
    x, y, z = _beam(100000, 5)
    n = len(x)
    rng = np.random.default_rng(6)
    u = [rng.normal(0, 1e3, n), rng.normal(0, 1e3, n), 1e6 + rng.normal(0, 1e3, n)]
    phi = _phi(7) * 100.
    q, m, dt = 1.602176565e-19, 39 * 1.660538921e-27, 1e-9
    g = [_t(a, device) for a in (x, y, z)] + [_t(a, device) for a in u]
    gi = torch.ones(n, dtype=torch.float64, device=device)
    E = [torch.zeros(n, dtype=torch.float64, device=device) for _ in range(3)]
    nlost = torch.zeros(1, dtype=torch.int64, device=device)
    N.call("wm_gather_push3d", _t(phi, device).data_ptr(), n, *[a.data_ptr() for a in g],
           gi.data_ptr(), *[a.data_ptr() for a in E], 0.0, *LO, DX, DY, DZ, NX, NY, NZ, q, m,
           dt, 0.02, 0.02, 0.08, nlost.data_ptr(), N.stream_of(device))
    xo, uo = [x.copy(), y.copy(), z.copy()], [a.copy() for a in u]
    eo = [np.zeros(n) for _ in range(3)]
    O.sete3d(phi, *xo, *eo, DX, DY, DZ, *LO)
    O.epush3d(*uo, *eo, q, m, 0.5 * (dt * 1 * 1.))
    O.epush3d(*uo, *eo, q, m, 0.5 * (dt * 1 * 1.))
    O.xpush3d(*xo, *uo, np.ones(n), dt)
    for a, b in zip(g + E, xo + uo + eo):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    X, Y, Z = xo
    out = ~((X >= LO[0]) & (X < 0.02) & (Y >= LO[1]) & (Y < 0.02) & (Z >= 0.) & (Z < 0.08))
    assert int(nlost.item()) == int(out.sum())


def test_getselfe_bitwise(device):
    from warp_amd import _native as N
    phi = _phi(8)
    se = torch.zeros((NZ + 1, NY + 1, NX + 1, 3), dtype=torch.float64, device=device)
    N.call("wm_getselfe3d", _t(phi, device).data_ptr(), NX, NY, NZ, se.data_ptr(), DX, DY, DZ,
           1, N.stream_of(device))
    np.testing.assert_array_equal(se.cpu().numpy(), O.getselfe3d(phi, DX, DY, DZ))


@pytest.mark.parametrize("bounds", [(1, 1, 1, 1, 1, 1), (2, 2, 2, 2, 2, 2), (0, 1, 2, 2, 1, 0)])
def test_rho_boundaries_bitwise(device, bounds):
    from warp_amd import _native as N
    r = np.random.default_rng(9).standard_normal((NZ + 1, NY + 1, NX + 1))
    rg = _t(r, device)
    b = (C.c_int64 * 6)(*bounds)
    N.call("wm_applyrhoboundaryconditions3d", NX, NY, NZ, rg.data_ptr(), b, N.stream_of(device))
    ro = r.copy()
    O.applyrhoboundaryconditions3d(ro, bounds)
    np.testing.assert_array_equal(rg.cpu().numpy(), ro)


def test_getzmmnt_matches_oracle(device):
    """Whole-beam moment sums on the GPU against the serial oracle (sums in
    another order: 1e-12 of each sum's scale), and the finalized
    emittances to 1e-9 (vz moments, which cancel ~1e6-fold, to 1e-5)."""
    from warp_amd import diagnostics as D
    from warp_amd.particles import ParticleGroup
    rng = np.random.default_rng(31)
    n = 300000
    x, y = rng.normal(1e-3, 4e-3, n), rng.normal(-2e-3, 3e-3, n)
    z = rng.normal(0.04, 0.01, n)
    ux, uy = 0.3e6 * x + rng.normal(0, 1e3, n), rng.normal(0, 1e3, n)
    uz = 1e6 + rng.normal(0, 1e3, n)
    pg = ParticleGroup(device)
    js = pg.addspecies(1.0, 1.0, 1.0)
    pg.addparticles(js, x[:n // 2], y[:n // 2], z[:n // 2], ux[:n // 2], uy[:n // 2], uz[:n // 2])
    js = pg.addspecies(1.0, 1.0, 1.0)
    pg.addparticles(js, x[n // 2:], y[n // 2:], z[n // 2:], ux[n // 2:], uy[n // 2:], uz[n // 2:])
    got = D.zmmnts0(pg)
    want = O.getzmmnt(x, y, z, ux, uy, uz, np.ones(n))
    scale = O.getzmmnt(np.abs(x), np.abs(y), np.abs(z), np.abs(ux), np.abs(uy), np.abs(uz),
                       np.ones(n))
    for k in range(33):
        assert abs(got[k] - want[k]) <= 1e-12 * scale[k], k
    a, b = D.finalize(got), D.finalize(want)
    for k in ("xrms", "yrms", "zrms", "xprms", "epsx", "epsy", "epsnx", "epsny", "epsr",
              "epsg", "epsh"):
        assert a[k] == pytest.approx(b[k], rel=1e-9), k
    for k in ("vzrms", "epsnz"):
        assert a[k] == pytest.approx(b[k], rel=1e-5), k


@pytest.mark.parametrize("center", ["halfv", "synchv"])
def test_padvnc3d_centers_bitwise(device, center):
    """wm_padvnc3d halfv (epush3d(dt/2), xpush3d(dt)) and synchv
    (epush3d(dt/2) only) equal the oracle's sequence bit for bit; synchv
    moves no particle and counts none lost."""
    from warp_amd import _native as N
    x, y, z = _beam(50000, 9)
    n = len(x)
    rng = np.random.default_rng(10)
    u = [rng.normal(0, 1e3, n), rng.normal(0, 1e3, n), 1e6 + rng.normal(0, 1e3, n)]
    phi = _phi(11) * 100.
    q, m, dt = 1.602176565e-19, 39 * 1.660538921e-27, 1e-9
    g = [_t(a, device) for a in (x, y, z)] + [_t(a, device) for a in u]
    gi = torch.ones(n, dtype=torch.float64, device=device)
    E = [torch.zeros(n, dtype=torch.float64, device=device) for _ in range(3)]
    nlost = torch.zeros(1, dtype=torch.int64, device=device)
    code = {"halfv": 1, "synchv": 2}[center]
    N.call("wm_padvnc3d", _t(phi, device).data_ptr(), n, *[a.data_ptr() for a in g],
           gi.data_ptr(), *[a.data_ptr() for a in E], 0.0, *LO, DX, DY, DZ, NX, NY, NZ, q, m,
           dt, 0.02, 0.02, 0.08, nlost.data_ptr(), code, N.stream_of(device))
    xo, uo = [x.copy(), y.copy(), z.copy()], [a.copy() for a in u]
    eo = [np.zeros(n) for _ in range(3)]
    O.sete3d(phi, *xo, *eo, DX, DY, DZ, *LO)
    O.epush3d(*uo, *eo, q, m, 0.5 * (dt * 1 * 1.))
    if center == "halfv":
        O.xpush3d(*xo, *uo, np.ones(n), dt)
    for a, b in zip(g + E, xo + uo + eo):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    if center == "synchv":
        assert int(nlost.item()) == 0


def test_padvnc3d_rejects_bad_center(device):
    from warp_amd import _native as N
    with pytest.raises(RuntimeError, match="center"):
        N.call("wm_padvnc3d", None, 0, None, None, None, None, None, None, None, None, None,
               None, 0.0, *LO, DX, DY, DZ, NX, NY, NZ, 1., 1., 1e-9, 0.02, 0.02, 0.08, None, 7,
               N.stream_of(device))


def test_sete_from_e_bitwise(device):
    """efetch=3 gather (sete_from_e_linear3d) on the GPU equals the oracle
    bit for bit, including particles on the upper faces and outside."""
    from warp_amd import _native as N
    phi = _phi(12)
    se = O.getselfe3d(phi, DX, DY, DZ)
    x, y, z = _beam(50000, 13)
    x = np.concatenate([x, [LO[0] + NX * DX, -0.03, 0.0]])
    y = np.concatenate([y, [0.0, 0.0, LO[1] + NY * DY]])
    z = np.concatenate([z, [0.04, 0.04, NZ * DZ]])
    n = len(x)
    e0 = [np.full(n, 0.25) for _ in range(3)]
    eo = [a.copy() for a in e0]
    O.sete_from_e3d(se, x, y, z, *eo, DX, DY, DZ, *LO)
    eg = [_t(a, device) for a in e0]
    sd, xd, yd, zd = _t(se, device), _t(x, device), _t(y, device), _t(z, device)
    N.call("wm_sete3d_from_e", sd.data_ptr(), n, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(),
           0.0, *LO, DX, DY, DZ, NX, NY, NZ, *[a.data_ptr() for a in eg], N.stream_of(device))
    for a, b in zip(eg, eo):
        np.testing.assert_array_equal(a.cpu().numpy(), b)


def test_padvnc3d_selfe_bitwise(device):
    """The fused fullv advance with efetch=3 equals E=0, sete_from_e,
    epush3d(dt/2) twice, xpush3d(dt) of the oracle."""
    from warp_amd import _native as N
    x, y, z = _beam(60000, 14)
    n = len(x)
    rng = np.random.default_rng(15)
    u = [rng.normal(0, 1e3, n), rng.normal(0, 1e3, n), 1e6 + rng.normal(0, 1e3, n)]
    se = O.getselfe3d(_phi(16) * 100., DX, DY, DZ)
    q, m, dt = 1.602176565e-19, 39 * 1.660538921e-27, 1e-9
    g = [_t(a, device) for a in (x, y, z)] + [_t(a, device) for a in u]
    gi = torch.ones(n, dtype=torch.float64, device=device)
    E = [torch.zeros(n, dtype=torch.float64, device=device) for _ in range(3)]
    sd = _t(se, device)
    N.call("wm_padvnc3d_selfe", sd.data_ptr(), n, *[a.data_ptr() for a in g],
           gi.data_ptr(), *[a.data_ptr() for a in E], 0.0, *LO, DX, DY, DZ, NX, NY, NZ, q, m,
           dt, 0.02, 0.02, 0.08, None, 0, N.stream_of(device))
    xo, uo = [x.copy(), y.copy(), z.copy()], [a.copy() for a in u]
    eo = [np.zeros(n) for _ in range(3)]
    O.sete_from_e3d(se, *xo, *eo, DX, DY, DZ, *LO)
    O.epush3d(*uo, *eo, q, m, 0.5 * (dt * 1 * 1.))
    O.epush3d(*uo, *eo, q, m, 0.5 * (dt * 1 * 1.))
    O.xpush3d(*xo, *uo, np.ones(n), dt)
    for a, b in zip(g + E, xo + uo + eo):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
