"""Pins of the FP64 oracle (CPU restatement of the reference) against the
reference's own known-answer test and against what the mathematics fixes:
closed forms, invariants and brute force on tiny inputs.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H


# ---------------------------------------------------------------------------
# multigrid3dsolve: the reference's known-answer test
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bounds", H.all_bounds())
@pytest.mark.parametrize("mgform", [1, 2])
def test_manufactured_kat(bounds, mgform):
    """warp_test/MutltiGrid3D_test.py: 24x22x18 on [-1,1]^3 for every
    boundary combination, mgtol 1e-12; the discrete solution is recovered
    to 1e-10 in fewer than mgmaxiters V-cycles.  mgform 2 (the residual-
    correction form, f3d_mgrid.F:336-372) must reach the same solution."""
    nx, ny, nz = 24, 22, 18
    phie, rho, phi, dx, dy, dz = H.manufactured(nx, ny, nz, bounds)
    rho0 = rho.copy()
    it, err = O.multigrid3dsolve(phi, rho, dx, dy, dz, bounds=bounds, mgtol=1e-12,
                                 mgform=mgform)
    assert it < 100 and err <= 1e-12
    # rho is scaled and restored in place (f3d_mgrid.F:224-230, 487-494)
    np.testing.assert_allclose(rho, rho0, rtol=1e-15, atol=0)
    t, p = phie.copy(), phi.copy()
    if 0 not in bounds:
        p -= p[1:-1, 1:-1, 1:-1].mean()
        t -= t[1:-1, 1:-1, 1:-1].mean()
    H.clearcorners(t)
    H.clearcorners(p)
    assert np.abs(p - t).max() < 1e-10


# ---------------------------------------------------------------------------
# sorhalfpass3d: Gauss-Seidel / SOR by hand on a tiny box
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mgparam", [1.0, 1.37])
def test_sor_by_hand(mgparam):
    """Red-black SOR written out point by point: points with
    (ix+iy+iz) % 2 == parity move to (1-w) phi + w * (Gauss-Seidel value
    of the 7-point equation -lap(phi) = rho/eps0).  The oracle takes rho
    pre-scaled by w/(2 eps0 sum(1/d^2)), as multigrid3dsolve hands it."""
    rng = np.random.default_rng(1)
    nx, ny, nz = 4, 5, 6
    dx, dy, dz = 0.3, 0.4, 0.5
    phi = rng.standard_normal((nz + 3, ny + 3, nx + 3))
    rho = rng.standard_normal((nz + 1, ny + 1, nx + 1))
    e0 = O.eps0()
    wx, wy, wz = 1 / dx ** 2, 1 / dy ** 2, 1 / dz ** 2
    scaled = rho * mgparam / (2 * e0 * (wx + wy + wz))
    for parity in (0, 1):
        ref = phi.copy()
        for iz in range(1, nz):
            for iy in range(1, ny):
                for ix in range(1, nx):
                    if (ix + iy + iz) % 2 != parity:
                        continue
                    P = lambda i, j, k: ref[k + 1, j + 1, i + 1]
                    gs = (rho[iz, iy, ix] / e0 + wx * (P(ix - 1, iy, iz) + P(ix + 1, iy, iz)) +
                          wy * (P(ix, iy - 1, iz) + P(ix, iy + 1, iz)) +
                          wz * (P(ix, iy, iz - 1) + P(ix, iy, iz + 1))) / (2 * (wx + wy + wz))
                    ref[iz + 1, iy + 1, ix + 1] = (1 - mgparam) * P(ix, iy, iz) + mgparam * gs
        O.sorhalfpass3d(parity, phi, scaled, dx, dy, dz, (0,) * 6, mgparam=mgparam)
        np.testing.assert_allclose(phi, ref, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------------------
# residual3d
# ---------------------------------------------------------------------------
def test_residual_zero_at_discrete_solution():
    """The residual of the manufactured discrete solution vanishes to
    rounding; the residual of phi = 0 is the scaled source."""
    nx, ny, nz = 12, 10, 8
    bounds = (0,) * 6
    phie, rho, _, dx, dy, dz = H.manufactured(nx, ny, nz, bounds)
    wx, wy, wz = 1 / dx ** 2, 1 / dy ** 2, 1 / dz ** 2
    mgparam = 1.2
    scaled = rho * mgparam / (2 * O.eps0() * (wx + wy + wz))
    r = O.residual3d(phie, scaled, dx, dy, dz, bounds, mgparam=mgparam)
    assert np.abs(r).max() < 1e-12 * np.abs(scaled).max()
    r0 = O.residual3d(np.zeros_like(phie), scaled, dx, dy, dz, bounds, mgparam=mgparam)
    np.testing.assert_allclose(r0[2:-2, 2:-2, 2:-2], scaled[1:-1, 1:-1, 1:-1], rtol=1e-15)
    # Dirichlet faces carry no residual
    assert not r0[1, :, :].any() and not r0[:, :, 1].any()


# ---------------------------------------------------------------------------
# restrict3d / expand3d
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,nc", [((16, 16, 16), (8, 8, 8)), ((16, 16, 32), (8, 8, 32))])
def test_restrict_constant_and_linear(n, nc):
    """Full weighting keeps a constant and reproduces a linear field at the
    coarse nodes, times ff; Dirichlet coarse faces are zero."""
    nx, ny, nz = n
    ncx, ncy, ncz = nc
    ff = 4.0
    for bounds in ((1,) * 6, (0,) * 6):
        res = np.full((nz + 3, ny + 3, nx + 3), 2.5)
        rc = O.restrict3d(res, ncx, ncy, ncz, ff, bounds)
        inner = rc if bounds[0] == 1 else rc[1:-1, 1:-1, 1:-1]
        np.testing.assert_allclose(inner, 2.5 * ff, rtol=1e-14)
        if bounds[0] == 0:
            assert not rc[0].any() and not rc[:, :, 0].any() and not rc[:, :, -1].any()
    # linear field, fine node i at x = i; coarse node I at x = I * nx/ncx
    k = np.arange(-1, nz + 2)[:, None, None]
    j = np.arange(-1, ny + 2)[None, :, None]
    i = np.arange(-1, nx + 2)[None, None, :]
    res = 0.3 * i - 0.7 * j + 1.1 * k + 5.
    rc = O.restrict3d(res, ncx, ncy, ncz, ff, (0,) * 6)
    K = np.arange(ncz + 1)[:, None, None] * (nz / ncz)
    J = np.arange(ncy + 1)[None, :, None] * (ny / ncy)
    I = np.arange(ncx + 1)[None, None, :] * (nx / ncx)
    want = ff * (0.3 * I - 0.7 * J + 1.1 * K + 5.)
    np.testing.assert_allclose(rc[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1], rtol=1e-13,
                               atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("n,nc", [((16, 16, 16), (8, 8, 8)), ((16, 16, 32), (8, 8, 32))])
def test_expand_linear_exact(n, nc):
    """Trilinear prolongation adds a linear coarse field exactly at every
    fine node off the Dirichlet faces."""
    nx, ny, nz = n
    ncx, ncy, ncz = nc
    K = np.arange(-1, ncz + 2)[:, None, None] * (nz / ncz)
    J = np.arange(-1, ncy + 2)[None, :, None] * (ny / ncy)
    I = np.arange(-1, ncx + 2)[None, None, :] * (nx / ncx)
    pc = 0.3 * I - 0.7 * J + 1.1 * K + 5.
    phi = np.full((nz + 3, ny + 3, nx + 3), 1.0)
    O.expand3d(phi, pc, (0,) * 6)
    k = np.arange(-1, nz + 2)[:, None, None]
    j = np.arange(-1, ny + 2)[None, :, None]
    i = np.arange(-1, nx + 2)[None, None, :]
    want = 1.0 + 0.3 * i - 0.7 * j + 1.1 * k + 5.
    np.testing.assert_allclose(phi[2:-2, 2:-2, 2:-2], want[2:-2, 2:-2, 2:-2], rtol=1e-13,
                               atol=1e-13 * np.abs(want).max())
    # Dirichlet faces untouched
    assert (phi[1] == 1.0).all() and (phi[:, :, -2] == 1.0).all()


def test_mglevels_of_the_configs():
    """getmglevels / getnextcoarselevel3d: z is not coarsened while
    dz > 4/3 min(d); coarsening stops below 4 cells."""
    lv = O.getmglevels(64, 64, 64, 0.04 / 64, 0.04 / 64, 0.08 / 64)
    assert [(l["nx"], l["ny"], l["nz"]) for l in lv] == [
        (64, 64, 64), (32, 32, 64), (16, 16, 32), (8, 8, 16), (4, 4, 8), (2, 2, 4)]
    assert len(O.getmglevels(128, 128, 256, 0.04 / 128, 0.04 / 128, 0.08 / 256)) == 7
    lv = O.getmglevels(256, 256, 512, 0.04 / 256, 0.04 / 256, 0.08 / 512)
    assert len(lv) == 8 and (lv[-1]["nx"], lv[-1]["nz"]) == (2, 4)
    # odd sizes round up to even
    assert O.getnextcoarselevel3d(13, 13, 13, 1., 1., 1.)[:3] == (8, 8, 8)


# ---------------------------------------------------------------------------
# boundary conditions
# ---------------------------------------------------------------------------
def test_applyboundaryconditions_rules():
    """Neumann mirrors u(-1) = u(1), periodic wraps u(-1) = u(n-1),
    Dirichlet with lwithdirichlet extrapolates u(-1) = 2u(0) - u(1)."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 6, 7, 8
    u = rng.standard_normal((nz + 3, ny + 3, nx + 3))
    v = u.copy()
    O.applyboundaryconditions3d(v, (1, 0, 2, 2, 0, 0), lwithdirichlet=True)
    c = slice(1, -1)
    np.testing.assert_array_equal(v[c, c, 0], v[c, c, 2])                     # x- Neumann
    np.testing.assert_allclose(v[c, c, -1], 2 * v[c, c, -2] - v[c, c, -3])    # x+ Dirichlet
    np.testing.assert_array_equal(v[c, 0, c], v[c, -3, c])                     # y periodic
    np.testing.assert_array_equal(v[c, -1, c], v[c, 2, c])
    np.testing.assert_allclose(v[0, c, c], 2 * v[1, c, c] - v[2, c, c])       # z- Dirichlet
    w = u.copy()
    O.applyboundaryconditions3d(w, (0,) * 6)
    np.testing.assert_array_equal(w, u)   # Dirichlet without lwithdirichlet: untouched


def test_rho_boundary_conditions():
    """applyrhoboundaryconditions3d (w3d.F:2660-2750, no guard cells):
    a Neumann face plane doubles, periodic faces sum into both planes."""
    rng = np.random.default_rng(4)
    rho = rng.standard_normal((6, 5, 4))
    r = rho.copy()
    O.applyrhoboundaryconditions3d(r, (1, 0, 2, 2, 0, 1))
    want = rho.copy()
    want[:, :, 0] *= 2
    s = want[:, 0, :] + want[:, -1, :]
    want[:, 0, :] = s
    want[:, -1, :] = s
    want[-1, :, :] *= 2
    np.testing.assert_allclose(r, want, rtol=1e-15)


# ---------------------------------------------------------------------------
# conductors: intercepts and subgrid distances against geometry
# ---------------------------------------------------------------------------
def _node(L, lev, p, idx):
    ix, iy, iz = idx
    return L["x0"] + ix * L["dx"][lev], L["y0"] + iy * L["dy"][lev], L["z0"] + iz * L["dz"][lev]


def _check_lists(lists, levels, g0, inside, on_surface):
    """Every subgrid distance < 1 lands on the surface (on_surface), every
    interior point is inside (inside), and on level 0 every node outside
    next to an inside node is a subgrid point (brute force)."""
    x0, y0, z0, dx, dy, dz = g0
    dirs = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))
    nsub = 0
    for p in (0, 1):
        ist = lists["sub%d_istart" % p]
        for lev in range(len(levels)):
            d = (dx * levels[lev]["lx"], dy * levels[lev]["ly"], dz * levels[lev]["lz"])
            for k in range(ist[lev], ist[lev + 1]):
                ix, iy, iz = lists["sub%d_indx" % p][k]
                pos = np.array([x0 + ix * d[0], y0 + iy * d[1], z0 + iz * d[2]])
                assert (ix + iy + iz) % 2 == p
                for m, (ax, sgn) in enumerate(dirs):
                    dl = lists["sub%d_dels" % p][k, m]
                    if dl < 1.0:
                        q = pos.copy()
                        q[ax] += sgn * dl * d[ax]
                        assert on_surface(q, d), (lev, (ix, iy, iz), m, dl)
                        nsub += 1
    assert nsub > 0
    ii = lists["int_istart"]
    for lev in range(len(levels)):
        d = (dx * levels[lev]["lx"], dy * levels[lev]["ly"], dz * levels[lev]["lz"])
        for k in range(ii[lev], ii[lev + 1]):
            ix, iy, iz = lists["int_indx"][k]
            assert inside(np.array([x0 + ix * d[0], y0 + iy * d[1], z0 + iz * d[2]]))


def _brute_adjacent(lists, n, g0, inside):
    nx, ny, nz = n
    x0, y0, z0, dx, dy, dz = g0
    sub = {tuple(v) for v in lists["int_indx"][:lists["int_istart"][1]]}  # pinned
    for p in (0, 1):
        e = lists["sub%d_istart" % p][1]
        sub |= {tuple(v) for v in lists["sub%d_indx" % p][:e]}
    ins = np.zeros((nz + 1, ny + 1, nx + 1), bool)
    for iz in range(nz + 1):
        for iy in range(ny + 1):
            for ix in range(nx + 1):
                ins[iz, iy, ix] = inside(np.array([x0 + ix * dx, y0 + iy * dy, z0 + iz * dz]))
    for iz in range(1, nz):
        for iy in range(1, ny):
            for ix in range(1, nx):
                if ins[iz, iy, ix]:
                    continue
                nb = (ins[iz, iy, ix - 1] or ins[iz, iy, ix + 1] or ins[iz, iy - 1, ix] or
                      ins[iz, iy + 1, ix] or ins[iz - 1, iy, ix] or ins[iz + 1, iy, ix])
                if nb:
                    assert (ix, iy, iz) in sub, (ix, iy, iz)


def _install(geoms, n, dom):
    co = O.install_conductors(geoms, *n, *dom)
    d = ((dom[1] - dom[0]) / n[0], (dom[3] - dom[2]) / n[1], (dom[5] - dom[4]) / n[2])
    O.checkconductors(co, *n, *d)
    levels = O.getmglevels(*n, *d)
    return co.lists(), levels, (dom[0], dom[2], dom[4]) + d


def test_zcylinder_intercepts_on_circle():
    """ZCylinder (f3d_conductors.F:2496-): x-line intercepts at
    xc +- sqrt(r^2 - (y-yc)^2), z intercepts on the end caps."""
    n, dom = (16, 16, 32), H.DOMAIN
    r, L, xc, yc, zc = 0.0061, 0.031, 0.0011, -0.0017, 0.0403
    lists, levels, g0 = _install([O.ZCylinder(r, L, 7., xcent=xc, ycent=yc, zcent=zc)], n, dom)

    def inside(q):
        return (q[0] - xc) ** 2 + (q[1] - yc) ** 2 < r * r and abs(q[2] - zc) < L / 2

    def on_surface(q, d):
        rr = np.hypot(q[0] - xc, q[1] - yc)
        on_side = abs(rr - r) < 1e-12 and abs(q[2] - zc) <= L / 2 + 1e-12
        on_cap = abs(abs(q[2] - zc) - L / 2) < 1e-12 and rr <= r + 1e-12
        return on_side or on_cap
    _check_lists(lists, levels, g0, inside, on_surface)
    _brute_adjacent(lists, n, g0, inside)
    assert (lists["int_volt"] == 7.).all()


def test_zcylinderout_and_annulus_intercepts():
    """ZCylinderOut (f3d_conductors.F:3158-3256) and the ZSrfrv annulus:
    intercepts on r = rout, r = rin, rmax and the planes z = zmin, zmax."""
    n, dom = (16, 16, 32), H.DOMAIN
    lists, levels, g0 = _install(H.pipe_and_annulus(O), n, dom)

    def inside(q, e=1e-12):
        # e > 0 counts nodes on the surface as inside (they are pinned)
        rr = np.hypot(q[0], q[1])
        if rr > H.PIPE_R - e:
            return True
        return H.ANN_RIN - e < rr < H.ANN_ROUT + e and H.ANN_Z0 - e < q[2] < H.ANN_Z1 + e

    def on_surface(q, d):
        rr = np.hypot(q[0], q[1])
        if abs(rr - H.PIPE_R) < 1e-12:
            return True
        if abs(rr - H.ANN_RIN) < 1e-12 and H.ANN_Z0 - 1e-12 <= q[2] <= H.ANN_Z1 + 1e-12:
            return True
        return (min(abs(q[2] - H.ANN_Z0), abs(q[2] - H.ANN_Z1)) < 1e-12 and
                H.ANN_RIN - 1e-12 <= rr <= H.ANN_ROUT + 1e-12)
    _check_lists(lists, levels, g0, inside, on_surface)
    # a node one cell from a surface node needs no subgrid stencil
    _brute_adjacent(lists, n, g0, lambda q: inside(q, -1e-12))


# ---------------------------------------------------------------------------
# conductors in the solve: closed forms
# ---------------------------------------------------------------------------
def _coax_error(n):
    nx = ny = n
    nz = 4
    dom = (-0.02, 0.02, -0.02, 0.02, 0.0, 0.01)
    a, b, V, xc, yc = 0.005, 0.018, 1000., 0.0003, -0.0002
    g = [O.ZCylinder(a, 0.1, V, xcent=xc, ycent=yc, zcent=0.005, condid=1),
         O.ZCylinderOut(b, 0.1, 0., xcent=xc, ycent=yc, zcent=0.005, condid=2)]
    co = O.install_conductors(g, nx, ny, nz, *dom)
    dx, dy, dz = 0.04 / nx, 0.04 / ny, 0.01 / nz
    phi = np.zeros((nz + 3, ny + 3, nx + 3))
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    it, err = O.multigrid3dsolve(phi, rho, dx, dy, dz, bounds=(0, 0, 0, 0, 1, 1),
                                 conductors=co, mgtol=1e-9, mgmaxiters=500)
    assert err <= 1e-9
    x = dom[0] + dx * np.arange(nx + 1)
    y = dom[2] + dy * np.arange(ny + 1)
    Y, X = np.meshgrid(y, x, indexing="ij")
    r = np.hypot(X - xc, Y - yc)
    exact = V * np.log(r / b) / np.log(a / b)
    m = (r > a) & (r < b)
    # uniform in z (Neumann ends)
    assert np.abs(phi[1:-1, 1:-1, 1:-1] - phi[3, 1:-1, 1:-1][None]).max() < 1e-8
    return np.abs(phi[3, 1:-1, 1:-1] - exact)[m].max()


def test_coax_second_order():
    """Shortley-Weller subgrid boundaries: the coaxial potential
    V ln(r/b)/ln(a/b) converges at second order under refinement."""
    e32, e64, e128 = _coax_error(32), _coax_error(64), _coax_error(128)
    assert e32 < 2e-3 * 1000.
    assert e32 / e64 > 3.5 and e64 / e128 > 3.5


@pytest.mark.parametrize("icndbndy", [1, 2])
@pytest.mark.parametrize("nz,z0", [(32, 0.05123), (16, 0.0313)])
def test_plane_at_fractional_offset_is_exactly_linear(nz, z0, icndbndy):
    """A conducting slab z >= z0 at V over phi(z=0) = 0 with Neumann x, y:
    the subgrid stencil reproduces phi = V z / z0 to round-off."""
    nx = ny = 4
    dom = (-0.02, 0.02, -0.02, 0.02, 0.0, 0.08)
    V = 100.
    co = O.install_conductors([O.ZCylinder(1.0, 0.3, V, zcent=z0 + 0.15, condid=1)],
                              nx, ny, nz, *dom)
    dz = 0.08 / nz
    phi = np.zeros((nz + 3, ny + 3, nx + 3))
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    it, err = O.multigrid3dsolve(phi, rho, 0.04 / nx, 0.04 / ny, dz, bounds=(1, 1, 1, 1, 0, 0),
                                 conductors=co, mgtol=1e-11, mgmaxiters=2000,
                                 icndbndy=icndbndy)
    z = dz * np.arange(nz + 1)
    k = z < z0
    want = (V * z[k] / z0)[:, None, None]
    assert np.abs(phi[1:-1, 1:-1, 1:-1][k] - want).max() < 1e-10


def test_reflection_symmetry():
    """Pipe + annulus with a source symmetric under x -> -x gives a
    potential symmetric to the solver tolerance."""
    nx, ny, nz = 32, 32, 64
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    dx, dy, dz = 0.04 / nx, 0.04 / ny, 0.08 / nz
    rng = np.random.default_rng(9)
    x = rng.normal(0, 0.004, 5000)
    x = np.concatenate([x, -x])
    y = np.tile(rng.normal(0, 0.004, 5000), 2)
    z = np.tile(0.04 + rng.normal(0, 0.01, 5000), 2)
    k = (np.abs(x) < 0.019) & (np.abs(y) < 0.019) & (z > 0.001) & (z < 0.079)
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O.setrho3d(rho, x[k], y[k], z[k], 1.6e-19, 1e6, dx, dy, dz, -0.02, -0.02, 0.)
    np.testing.assert_allclose(rho, rho[:, :, ::-1], rtol=1e-12, atol=1e-12 * rho.max())
    phi = np.zeros((nz + 3, ny + 3, nx + 3))
    it, err = O.multigrid3dsolve(phi, rho, dx, dy, dz, conductors=co, mgtol=1e-8,
                                 mgmaxiters=300)
    assert err <= 1e-8
    assert np.abs(phi - phi[:, :, ::-1]).max() < 1e-6


# ---------------------------------------------------------------------------
# particles
# ---------------------------------------------------------------------------
def _cic_brute(rho, x, y, z, q, w, dx, dy, dz, x0, y0, z0):
    for xp, yp, zp in zip(x, y, z):
        fx, fy, fz = (xp - x0) / dx, (yp - y0) / dy, (zp - z0) / dz
        i, j, k = int(np.floor(fx)), int(np.floor(fy)), int(np.floor(fz))
        u, v, t = fx - i, fy - j, fz - k
        for a, wa in ((0, 1 - u), (1, u)):
            for b, wb in ((0, 1 - v), (1, v)):
                for c, wc in ((0, 1 - t), (1, t)):
                    rho[k + c, j + b, i + a] += q * w * wa * wb * wc / (dx * dy * dz)


def test_deposit_brute_force_and_charge():
    rng = np.random.default_rng(5)
    nx, ny, nz = 4, 5, 6
    dx, dy, dz = 0.1, 0.2, 0.3
    x0, y0, z0 = -0.2, 0.1, 1.0
    npart = 200
    x = x0 + rng.uniform(0, nx * dx, npart) * 0.999
    y = y0 + rng.uniform(0, ny * dy, npart) * 0.999
    z = z0 + rng.uniform(0, nz * dz, npart) * 0.999
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O.setrho3d(rho, x, y, z, 2.0, 3.0, dx, dy, dz, x0, y0, z0)
    ref = np.zeros_like(rho)
    _cic_brute(ref, x, y, z, 2.0, 3.0, dx, dy, dz, x0, y0, z0)
    np.testing.assert_allclose(rho, ref, rtol=1e-12, atol=1e-12 * ref.max())
    assert abs(rho.sum() * dx * dy * dz - 2.0 * 3.0 * npart) < 1e-10 * npart
    # a particle on a node puts all its charge there (dyadic mesh: exact)
    dx, dy, dz, x0, y0, z0 = 0.125, 0.25, 0.5, -0.5, 0.25, 1.0
    r1 = np.zeros_like(rho)
    O.setrho3d(r1, [x0 + 2 * dx], [y0 + 3 * dy], [z0 + 1 * dz], 1.0, 1.0, dx, dy, dz, x0, y0, z0)
    assert r1[1, 3, 2] == pytest.approx(1. / (dx * dy * dz), rel=1e-12)
    assert np.count_nonzero(r1) == 1


def test_gather_linear_field_exact():
    """sete_from_phi_linear3d: a linear phi gives E = -grad phi at any
    position; a particle on a node gets the centred difference there."""
    rng = np.random.default_rng(6)
    nx, ny, nz = 6, 7, 8
    dx, dy, dz = 0.1, 0.2, 0.15
    x0, y0, z0 = 0.3, -0.1, 0.0
    K = np.arange(-1, nz + 2)[:, None, None] * dz + z0
    J = np.arange(-1, ny + 2)[None, :, None] * dy + y0
    I = np.arange(-1, nx + 2)[None, None, :] * dx + x0
    phi = 2.0 * I - 3.0 * J + 0.5 * K + 1.0
    n = 100
    x = x0 + rng.uniform(0, nx * dx * 0.999, n)
    y = y0 + rng.uniform(0, ny * dy * 0.999, n)
    z = z0 + rng.uniform(0, nz * dz * 0.999, n)
    ex, ey, ez = np.zeros(n), np.zeros(n), np.zeros(n)
    O.sete3d(phi, x, y, z, ex, ey, ez, dx, dy, dz, x0, y0, z0)
    np.testing.assert_allclose(ex, -2.0, rtol=1e-12)
    np.testing.assert_allclose(ey, 3.0, rtol=1e-12)
    np.testing.assert_allclose(ez, -0.5, rtol=1e-12)
    # accumulates into E (w3d.F:4103-4105)
    O.sete3d(phi, x, y, z, ex, ey, ez, dx, dy, dz, x0, y0, z0)
    np.testing.assert_allclose(ex, -4.0, rtol=1e-12)
    p2 = rng.standard_normal(phi.shape)
    e = [np.zeros(1) for _ in range(3)]
    O.sete3d(p2, [x0 + 2 * dx], [y0 + 3 * dy], [z0 + 4 * dz], *e, dx, dy, dz, x0, y0, z0)
    P = lambda i, j, k: p2[k + 1, j + 1, i + 1]
    assert e[0][0] == pytest.approx((P(1, 3, 4) - P(3, 3, 4)) / (2 * dx), rel=1e-12)
    assert e[1][0] == pytest.approx((P(2, 2, 4) - P(2, 4, 4)) / (2 * dy), rel=1e-12)
    assert e[2][0] == pytest.approx((P(2, 3, 3) - P(2, 3, 5)) / (2 * dz), rel=1e-12)


def test_leapfrog_constant_field_exact():
    """fullv ordering (half kick, half kick, drift) in a constant field:
    after k steps x = x0 + k u0 dt + a dt^2 k (k+1)/2, u = u0 + k a dt."""
    q, m, dt = 1.6e-19, 6.5e-26, 1e-9
    E = np.array([1.0e4, -2.0e3, 5.0e2])
    a = q * E / m
    x = np.array([0.001]); y = np.array([-0.002]); z = np.array([0.03])
    ux = np.array([10.]); uy = np.array([-20.]); uz = np.array([1e6])
    X0 = np.array([x[0], y[0], z[0]])
    U0 = np.array([ux[0], uy[0], uz[0]])
    g = np.ones(1)
    kmax = 50
    for _ in range(kmax):
        ex, ey, ez = (np.full(1, E[0]), np.full(1, E[1]), np.full(1, E[2]))
        O.epush3d(ux, uy, uz, ex, ey, ez, q, m, 0.5 * dt)
        O.epush3d(ux, uy, uz, ex, ey, ez, q, m, 0.5 * dt)
        O.xpush3d(x, y, z, ux, uy, uz, g, dt)
    want_u = U0 + kmax * a * dt
    want_x = X0 + kmax * U0 * dt + a * dt * dt * kmax * (kmax + 1) / 2
    np.testing.assert_allclose([ux[0], uy[0], uz[0]], want_u, rtol=1e-12)
    np.testing.assert_allclose([x[0], y[0], z[0]], want_x, rtol=1e-12, atol=1e-15)


def test_getselfe_linear():
    nx, ny, nz = 5, 6, 7
    dx, dy, dz = 0.2, 0.1, 0.3
    K = np.arange(-1, nz + 2)[:, None, None] * dz
    J = np.arange(-1, ny + 2)[None, :, None] * dy
    I = np.arange(-1, nx + 2)[None, None, :] * dx
    phi = -1.5 * I + 2.5 * J + 4.0 * K
    se = O.getselfe3d(phi, dx, dy, dz)
    np.testing.assert_allclose(se[..., 0], 1.5, rtol=1e-12)
    np.testing.assert_allclose(se[..., 1], -2.5, rtol=1e-12)
    np.testing.assert_allclose(se[..., 2], -4.0, rtol=1e-12)


def test_sete_from_e_linear_field_exact():
    """Trilinear interpolation reproduces a field linear in x, y, z, and a
    particle on a node gets that node's value (sete_from_e_linear3d)."""
    nx, ny, nz, dx, dy, dz = 8, 6, 10, 0.1, 0.2, 0.3
    lo = (-0.4, -0.6, 0.1)
    Z, Y, X = np.meshgrid(lo[2] + np.arange(nz + 1) * dz, lo[1] + np.arange(ny + 1) * dy,
                          lo[0] + np.arange(nx + 1) * dx, indexing="ij")
    coef = np.array([[1., 2., -3., 0.5], [0.3, -1., 4., 2.], [-2., 0.1, 0.2, -0.7]])
    se = np.stack([c[0] + c[1] * X + c[2] * Y + c[3] * Z for c in coef], -1)
    rng = np.random.default_rng(3)
    n = 2000
    x = lo[0] + rng.uniform(0, nx * dx, n)
    y = lo[1] + rng.uniform(0, ny * dy, n)
    z = lo[2] + rng.uniform(0, nz * dz, n)
    e = [np.zeros(n) for _ in range(3)]
    O.sete_from_e3d(se, x, y, z, *e, dx, dy, dz, *lo)
    for c, ec in zip(coef, e):
        np.testing.assert_allclose(ec, c[0] + c[1] * x + c[2] * y + c[3] * z, atol=1e-12)
    # nodes on a dyadic mesh (scaled positions exact), including the upper
    # faces, where the weight-zero neighbour lies beyond the mesh
    bx, by, bz = 0.125, 0.25, 0.5
    ii = np.array([0, 3, nx, nx]); jj = np.array([0, 2, ny, 0]); kk = np.array([0, 5, nz, nz])
    xn, yn, zn = ii * bx, jj * by, kk * bz
    e = [np.zeros(4) for _ in range(3)]
    O.sete_from_e3d(se, xn, yn, zn, *e, bx, by, bz, 0., 0., 0.)
    for c in range(3):
        assert np.array_equal(e[c], se[kk, jj, ii, c])
    # accumulates, and skips particles outside [0, n]
    e = [np.ones(2) for _ in range(3)]
    O.sete_from_e3d(se, np.array([-0.01, xn[1]]), yn[:2], zn[:2], *e, bx, by, bz, 0., 0., 0.)
    assert e[0][0] == 1. and e[0][1] == 1. + se[kk[1], jj[1], ii[1], 0]


def test_efetch3_of_linear_phi_is_constant():
    """getselfe3d then sete_from_e_linear3d of a linear phi gives -grad phi
    at any position, as the phi gather (efetch 1) does."""
    nx, ny, nz, dx, dy, dz = 5, 6, 7, 0.2, 0.1, 0.3
    K = np.arange(-1, nz + 2)[:, None, None] * dz
    J = np.arange(-1, ny + 2)[None, :, None] * dy
    I = np.arange(-1, nx + 2)[None, None, :] * dx
    phi = -1.5 * I + 2.5 * J + 4.0 * K
    se = O.getselfe3d(phi, dx, dy, dz)
    rng = np.random.default_rng(4)
    x, y, z = rng.uniform(0, nx * dx, 300), rng.uniform(0, ny * dy, 300), rng.uniform(0, nz * dz, 300)
    e = [np.zeros(300) for _ in range(3)]
    O.sete_from_e3d(se, x, y, z, *e, dx, dy, dz, 0., 0., 0.)
    np.testing.assert_allclose(e[0], 1.5, rtol=1e-12)
    np.testing.assert_allclose(e[1], -2.5, rtol=1e-12)
    np.testing.assert_allclose(e[2], -4.0, rtol=1e-12)


# ---------------------------------------------------------------------------
# beam moments (getzmmnt) and emittances
# ---------------------------------------------------------------------------
def _moment_table(x, y, z, ux, uy, uz):
    """zmmnts0(k) named as top.F:1366-1399 reads them back (k = 1..33)."""
    xp, yp = ux / uz, uy / uz
    return {1: len(x), 2: x.sum(), 3: y.sum(), 4: z.sum(), 5: xp.sum(), 6: yp.sum(),
            7: ux.sum(), 8: uy.sum(), 9: uz.sum(), 10: (x * y).sum(), 11: (xp * yp).sum(),
            12: (x * x).sum(), 13: (y * y).sum(), 14: (z * z).sum(), 15: (xp * xp).sum(),
            16: (yp * yp).sum(), 17: (ux * ux).sum(), 18: (uy * uy).sum(), 19: (uz * uz).sum(),
            20: (x * xp).sum(), 21: (y * yp).sum(), 22: (z * uz).sum(), 23: (x * yp).sum(),
            24: (y * xp).sum(), 25: (x * uz).sum(), 26: (y * uz).sum(), 27: (ux * uz).sum(),
            28: (uy * uz).sum(), 29: (x * ux).sum(), 30: (y * uy).sum(), 31: (x * uy).sum(),
            32: (y * ux).sum(), 33: (ux * uy).sum()}


def test_getzmmnt_sums_named():
    """Each zmmnts0(k) is the sum the finalization names it by."""
    rng = np.random.default_rng(21)
    n = 5000
    x, y, z = rng.normal(1e-3, 4e-3, n), rng.normal(-2e-3, 3e-3, n), rng.normal(0.04, 0.01, n)
    ux, uy, uz = rng.normal(10, 1e3, n), rng.normal(-5, 1e3, n), 1e6 + rng.normal(0, 1e3, n)
    mm = O.getzmmnt(x, y, z, ux, uy, uz, np.ones(n))
    want = _moment_table(x, y, z, ux, uy, uz)
    for k, v in want.items():
        assert mm[k - 1] == pytest.approx(v, rel=1e-11, abs=1e-300), k
    # the gaminv branch with gaminv = 1 gives the same sums
    np.testing.assert_allclose(O.getzmmnt(x, y, z, ux, uy, uz, np.ones(n), True), mm, rtol=1e-15)


def test_emittance_closed_forms():
    """An ellipse x = a cos t, x' = b sin t (t uniform over a period) has
    epsx = 2ab; a beam with x' = k x has zero emittance."""
    from warp_amd.diagnostics import finalize
    n = 4096
    t = 2 * np.pi * (np.arange(n) + 0.5) / n
    a, b, uz = 3e-3, 2e-3, 1e6
    x, ux = a * np.cos(t), b * uz * np.sin(t)
    y, uy = 2 * a * np.sin(t), 0.5 * b * uz * np.cos(t)
    z, uzv = np.full(n, 0.04), np.full(n, uz)
    d = finalize(O.getzmmnt(x, y, z, ux, uy, uzv, np.ones(n)))
    assert d["epsx"] == pytest.approx(2 * a * b, rel=1e-12)
    assert d["epsy"] == pytest.approx(2 * (2 * a) * (0.5 * b), rel=1e-12)
    assert d["xrms"] == pytest.approx(a / np.sqrt(2), rel=1e-12)
    d2 = finalize(O.getzmmnt(x, y, z, 0.7 * uz * x, uy, uzv, np.ones(n)))
    assert d2["epsx"] < 1e-6 * d["epsx"]   # sqrt of a rounding-level difference
    # normalized: epsnx = 4 sqrt(<x^2><vx^2> - <x vx>^2) / c * 1e6 = epsx * uz / c * 1e6
    assert d["epsnx"] == pytest.approx(d["epsx"] * uz / 2.99792458e8 * 1e6, rel=1e-12)


# ---------------------------------------------------------------------------
# conductor-corrected E (lcorrectede)
# ---------------------------------------------------------------------------
SLAB_N3, SLAB_D3 = (9, 7, 8), (0.1, 0.2, 0.15)
SLAB_COEF = (-7.5, 4.0, 9.0)


def slab_case(axis, side, frac, efetch, ic=4, V=3.0, n=3000, seed=0):
    """Corrected E of random particles near a slab (oracle); returns
    (CondField, phi, selfe, positions, E) for the checks."""
    n3, d3 = SLAB_N3, SLAB_D3
    d = H.slab_lists(n3, axis, side, ic, frac, V)
    phi = H.slab_phi(n3, d3, axis, side, ic, frac, V, SLAB_COEF)
    cf = O.CondField(d, *n3)
    cf.efield(phi, *d3)
    selfe = None
    if efetch == 3:
        selfe = O.getselfe3d(phi, *d3)
        cf.fixefield(selfe)
    rng = np.random.default_rng(seed + 10 * axis + (side > 0))
    pos = [rng.uniform(0, n3[a] * d3[a], n) for a in range(3)]
    e = [np.zeros(n) for _ in range(3)]
    O.sete3d_conductor(cf, phi, selfe, *pos, *e, *d3, 0., 0., 0., efetch=efetch)
    return cf, phi, selfe, pos, e


def slab_expected(axis, side, frac, pos, ic=4):
    """The field the corrected gather must give near the slab of
    slab_case: with phi = V + (s - s_c) f(t) on the vacuum nodes, the
    centred differences there and the one-sided subgrid E (and its
    extrapolation into the conductor) are the exact nodal field
    E_s = -f(t_j), E_t = -(s - s_c) f'(t_j), E_r = 0, so the gather is the
    interpolation of the exact nodal field: linear in t for E_s; for E_t,
    bilinear in (s, t) in vacuum cells and, in the cell cut by the
    surface, the vacuum face's value -(s_ic - s_c) f'(t).  Cells whose
    nodes are all covered get E = 0."""
    n3, d3 = SLAB_N3, SLAB_D3
    ta, ra = (axis + 1) % 3, (axis + 2) % 3
    g, h, q = SLAB_COEF
    s, t = pos[axis], pos[ta]
    si = np.floor(s / d3[axis]).astype(int)
    tj = np.floor(t / d3[ta]).astype(int)
    v = t / d3[ta] - tj
    t0, t1 = tj * d3[ta], (tj + 1) * d3[ta]
    sc = H.slab_surface(axis, side, ic, frac, d3)
    fp = lambda tt: h + 2 * q * tt
    es = -((1 - v) * H.slab_f(t0, SLAB_COEF) + v * H.slab_f(t1, SLAB_COEF))
    et = -(s - sc) * ((1 - v) * fp(t0) + v * fp(t1))
    if side < 0:
        cut, inside = si == ic - 1, si < ic - 1
    else:
        cut, inside = si == ic, si > ic
    et = np.where(cut, -(ic * d3[axis] - sc) * ((1 - v) * fp(t0) + v * fp(t1)), et)
    es = np.where(inside, 0., es)
    et = np.where(inside, 0., et)
    out = [None] * 3
    out[axis], out[ta], out[ra] = es, et, np.zeros_like(es)
    return out, cut, inside


@pytest.mark.parametrize("efetch", [1, 3])
@pytest.mark.parametrize("frac", [0.4, 0.85, 1.0])
@pytest.mark.parametrize("axis,side", [(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)])
def test_correctede_slab_exact(axis, side, frac, efetch):
    """A slab at V normal to each axis, on either side, with
    phi = V + (s - s_c) f(t) beyond it: the one-sided subgrid E (efield0)
    and its extrapolation into the conductor (efieldd) are the exact field
    at the subgrid nodes, and every particle gets the interpolation of the
    exact nodal field, including in the cell cut by the surface, where
    centred differences of phi are wrong; cells with only covered nodes get
    E = 0."""
    cf, phi, selfe, pos, e = slab_case(axis, side, frac, efetch)
    want, cut, inside = slab_expected(axis, side, frac, pos)
    ta = (axis + 1) % 3
    k = 2 * axis + (0 if side < 0 else 1)
    for d in cf.sub:
        tn = d["indx"][:, ta] * SLAB_D3[ta]
        np.testing.assert_allclose(d["efield0"][:, axis], -H.slab_f(tn, SLAB_COEF), rtol=1e-12)
        np.testing.assert_allclose(d["efieldd"][:, k], -H.slab_f(tn, SLAB_COEF), rtol=1e-12)
        others = [c for c in range(6) if c != k]
        assert np.all(d["efieldd"][:, others] == 0.)   # only one side is covered
    assert cut.sum() > 100 and inside.sum() > 100
    for c in range(3):
        scale = np.abs(want[c]).max() + 1e-300
        np.testing.assert_allclose(e[c], want[c], rtol=0, atol=1e-12 * scale)
        assert np.all(e[c][inside] == 0.)


def test_correctede_fixes_what_plain_gather_gets_wrong():
    """The same slab through the plain gather: the cell cut by the surface
    gets the wrong field (the check above is not vacuous)."""
    axis, side, frac, ic = 0, -1, 0.4, 4
    phi = H.slab_phi(SLAB_N3, SLAB_D3, axis, side, ic, frac, 3.0, SLAB_COEF)
    _, _, _, pos, _ = slab_case(axis, side, frac, 1)
    want, cut, _ = slab_expected(axis, side, frac, pos)
    e = [np.zeros(int(cut.sum())) for _ in range(3)]
    inmesh = [p[cut] for p in pos]
    O.sete3d(phi, *inmesh, *e, *SLAB_D3, 0., 0., 0.)
    assert np.abs(e[0] - want[0][cut]).max() > 0.1 * np.abs(want[0]).max()


def test_correctede_vacuum_is_the_plain_gather():
    """With no conductors every cell takes the vacuum path: efetch 3 is
    bit-for-bit sete_from_e_linear3d; efetch 1 differs from
    sete_from_phi_linear3d only by where 1/(2 dx) multiplies (per term in
    sete3dwithconductor, once per sum in sete3d)."""
    nx, ny, nz, dx, dy, dz = 6, 5, 7, 0.1, 0.12, 0.09
    rng = np.random.default_rng(8)
    phi = rng.standard_normal((nz + 3, ny + 3, nx + 3))
    empty = dict(levels=1, int_indx=np.zeros((0, 3), np.int64), int_volt=np.zeros(0),
                 int_istart=np.zeros(2, np.int64))
    for p in range(2):
        empty["sub%d_indx" % p] = np.zeros((0, 3), np.int64)
        empty["sub%d_dels" % p] = np.zeros((0, 6))
        empty["sub%d_volt" % p] = np.zeros((0, 6))
        empty["sub%d_istart" % p] = np.zeros(2, np.int64)
    cf = O.CondField(empty, nx, ny, nz)
    assert not cf.icgrid.any()
    n = 500
    x = rng.uniform(0, (nx - 1e-9) * dx, n)
    y = rng.uniform(0, (ny - 1e-9) * dy, n)
    z = rng.uniform(0, (nz - 1e-9) * dz, n)
    se = O.getselfe3d(phi, dx, dy, dz)
    a = [np.zeros(n) for _ in range(3)]
    b = [np.zeros(n) for _ in range(3)]
    O.sete3d_conductor(cf, phi, se, x, y, z, *a, dx, dy, dz, 0., 0., 0., efetch=3)
    O.sete_from_e3d(se, x, y, z, *b, dx, dy, dz, 0., 0., 0.)
    for c in range(3):
        assert np.array_equal(a[c], b[c])
    a = [np.zeros(n) for _ in range(3)]
    b = [np.zeros(n) for _ in range(3)]
    O.sete3d_conductor(cf, phi, None, x, y, z, *a, dx, dy, dz, 0., 0., 0., efetch=1)
    O.sete3d(phi, x, y, z, *b, dx, dy, dz, 0., 0., 0.)
    for c in range(3):
        np.testing.assert_allclose(a[c], b[c], rtol=1e-13, atol=1e-13 * np.abs(b[c]).max())


def test_icgrid_rules():
    """setupconductorfielddata: interior points marked above every record
    index, even records +(i+1), odd -(i+1), later lists overwrite."""
    d = H.slab_lists((6, 3, 3), 0, -1, 2, 0.5, 1.0)
    # put the first even record's node into the interior list too: the even
    # list, written later, wins
    d["int_indx"] = np.vstack([d["int_indx"], d["sub0_indx"][:1]])
    d["int_volt"] = np.append(d["int_volt"], 1.0)
    d["int_istart"] = np.array([0, len(d["int_volt"])], np.int64)
    cf = O.CondField(d, 6, 3, 3)
    ic = cf.icgrid
    assert np.all(ic[:, :, :2] == O.OC_INTERIOR)
    for p, sgn in ((0, 1), (1, -1)):
        for i, (x, y, z) in enumerate(d["sub%d_indx" % p]):
            assert ic[z, y, x] == sgn * (i + 1)
    assert np.all(ic[:, :, 3:] == 0)


def test_correctede_empty_parity_list_skips_subgrid_field():
    """getefieldatconductorsubgrid returns at once when either parity list
    is empty (f3d_mgrid.F:4659-4661): the other list's E stays zero."""
    d = H.slab_lists((6, 0, 0), 0, -1, 2, 0.5, 1.0)   # a single record, even
    assert len(d["sub0_indx"]) == 1 and len(d["sub1_indx"]) == 0
    cf = O.CondField(d, 6, 0, 0)
    cf.efield(H.slab_phi((6, 0, 0), (0.1,) * 3, 0, -1, 2, 0.5, 1.0, (2., 0., 0.)), 0.1, 0.1, 0.1)
    assert not cf.sub[0]["efield0"].any()


NEUMANN_N3, NEUMANN_D3, NEUMANN_GH = (12, 8, 16), (0.1, 0.12, 0.08), (2.0, 25.0)


@pytest.mark.parametrize("kind", ["m", "p", "mp", "dm"])
def test_neumann_sheet_closed_form(kind):
    """Neumann subgrid points (dels in (-1, 0]): a field constant in x,
    phi = g + h z, has zero normal derivative on a sheet normal to x, so it
    solves the subgrid equations of handlesubgrid2's Neumann branches (and
    the Dirichlet branch at the voltage phi takes there) and condbndymgint's
    mirror leaves it unchanged; from a perturbed start the solve must
    return it.  A wrong Neumann coefficient (pp/(0.5+adel), the denominator
    term 2 adel/(adel+0.5)) or mirror would make the solution vary in x.

    Dirichlet below and Neumann above ("dm"): the reference's denominator
    term (1 - 2 dels adel - dels)/(dels (adel + dels/2)) (f3d_mgrid.F:2702-
    2703) against pm = volt/(dels (adel + dels/2)) makes the field constant
    in x the solution when the surface voltage is phi (1 - dels + dels^2)
    rather than phi (the term 1 - 2 dels adel - dels^2 would give phi); the
    test sets that voltage, so it holds the restatement to the reference's
    form of this branch."""
    n3, d3 = NEUMANN_N3, NEUMANN_D3
    g, h = NEUMANN_GH
    dl = 0.3
    vf = 1. - dl + dl * dl if kind == "dm" else 1.
    phi, exact = H.linear_z_case(n3, d3, g, h, seed=3)
    d = H.neumann_lists(n3, 5, kind, dl, lambda ix, iy, iz: (g + h * iz * d3[2]) * vf)
    co = H.oracle_conductors(d)
    rho = np.zeros((n3[2] + 1, n3[1] + 1, n3[0] + 1))
    it, err = O.multigrid3dsolve(phi, rho, *d3, conductors=co, mgtol=1e-11, mgmaxiters=200)
    assert err <= 1e-11 and it < 200
    inner = (slice(1, n3[2] + 2), slice(1, n3[1] + 2), slice(1, n3[0] + 2))
    assert np.abs(phi[inner] - exact[inner]).max() < 1e-8
    # the same lists with the sheet's dels made Dirichlet at voltage 0
    # (a sign error in the branch selection) give a different field
    d0 = dict(d)
    for p in range(2):
        d0["sub%d_dels" % p] = np.abs(d["sub%d_dels" % p])
        d0["sub%d_volt" % p] = np.zeros_like(d["sub%d_volt" % p])
    phi0, _ = H.linear_z_case(n3, d3, g, h, seed=3)
    O.multigrid3dsolve(phi0, rho, *d3, conductors=H.oracle_conductors(d0), mgtol=1e-11,
                       mgmaxiters=200)
    assert np.abs(phi0[inner] - exact[inner]).max() > 1e-2


def test_fixefield_mesh_aligned_face():
    """A conductor face on a mesh plane (no subgrid points): the centred
    grid E at the face node is half the vacuum field; doubling it at the
    interior points restores -g there and leaves zero deeper inside."""
    nx, ny, nz, dx, V, g, ic = 8, 3, 3, 0.1, 2.0, 5.0, 3
    d = H.slab_lists((nx, ny, nz), 0, -1, ic + 1, 1.0, V)     # covered nodes ix <= ic
    x = (np.arange(-1, nx + 2) - ic) * dx
    phi = np.ascontiguousarray(np.broadcast_to(np.where(x <= 0, V, V + g * x),
                                               (nz + 3, ny + 3, nx + 3)))
    se = O.getselfe3d(phi, 0.1, 0.1, 0.1)
    np.testing.assert_allclose(se[:, :, ic, 0], -g / 2, rtol=1e-12)
    cf = O.CondField(d, nx, ny, nz)
    cf.fixefield(se)
    np.testing.assert_allclose(se[:, :, ic, 0], -g, rtol=1e-12)
    assert np.all(se[:, :, :ic, 0] == 0.)
    np.testing.assert_allclose(se[:, :, ic + 1:nx, 0], -g, rtol=1e-12)


# ---------------------------------------------------------------------------
# Field diagnostics: line charge (gtlchg3dfromrho) and electrostatic energy
# (getese3dfromrhophi)
# ---------------------------------------------------------------------------
def _with_guards(interior, fill=1.0e30):
    """phi with one guard cell around interior, the guards holding a value
    that would dominate any sum they entered."""
    return np.pad(interior, 1, constant_values=fill)


def test_getese_constant_fields_closed_form():
    """rho = 1/4 and phi = 2 on every node: ese = dx*dy*dz/2 * 1/2 * (number
    of nodes), exact in binary; the 1e30 guard cells of phi must not enter."""
    nx, ny, nz = 6, 4, 5
    dx, dy, dz = 0.5, 0.25, 0.125
    rho = np.full((nz + 1, ny + 1, nx + 1), 0.25)
    phi = _with_guards(np.full((nz + 1, ny + 1, nx + 1), 2.0))
    e = O.getese3d(rho, phi, dx, dy, dz)
    assert e == 0.5 * dx * dy * dz * 0.5 * (nx + 1) * (ny + 1) * (nz + 1)


def test_getese_is_half_the_integral_of_rho_phi():
    rng = np.random.default_rng(11)
    nx, ny, nz = 9, 7, 12
    dx, dy, dz = 0.1, 0.2, 0.05
    rho = rng.standard_normal((nz + 1, ny + 1, nx + 1))
    inner = rng.standard_normal((nz + 1, ny + 1, nx + 1))
    e = O.getese3d(rho, _with_guards(inner), dx, dy, dz)
    assert e == pytest.approx(0.5 * dx * dy * dz * float(np.sum(rho * inner)), rel=1e-13)


def _mirror(a, axes):
    """The full mesh of a field stored only for index >= 0 along axes
    (numpy axes of a (z, y, x) array): reflected through index 0."""
    for ax in axes:
        a = np.concatenate([np.flip(a, axis=ax).take(range(a.shape[ax] - 1), axis=ax), a],
                           axis=ax)
    return a


@pytest.mark.parametrize("sym", ["l2symtry", "l4symtry"])
def test_symmetry_factors_equal_the_reflected_full_mesh(sym):
    """l4symtry stores the quadrant x, y >= 0: each plane's charge and the
    energy equal those of the full mesh made by reflecting it through x = 0
    and y = 0.  l2symtry: the plane charge reflects through y = 0
    (dw3d.F:262-265) and the energy through x = 0 (the reference doubles the
    columns ix >= 1, dw3d.F:752-756)."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 5, 4, 6
    dx, dy, dz = 0.1, 0.2, 0.05
    rho = rng.standard_normal((nz + 1, ny + 1, nx + 1))
    inner = rng.standard_normal((nz + 1, ny + 1, nx + 1))
    kw = {sym: True}
    if sym == "l4symtry":
        ax_lc, ax_e = (1, 2), (1, 2)
    else:
        ax_lc, ax_e = (1,), (2,)
    full_planes = O.plane_sums(_mirror(rho, ax_lc))
    np.testing.assert_allclose(O.plane_sums(rho, **kw), full_planes, rtol=1e-13, atol=0)
    ef = O.getese3d(_mirror(rho, ax_e), _with_guards(_mirror(inner, ax_e)), dx, dy, dz)
    assert O.getese3d(rho, _with_guards(inner), dx, dy, dz, **kw) == pytest.approx(ef, rel=1e-13)


def test_line_charge_integrates_to_the_deposited_charge():
    """CIC charge of random particles (setrho3d) summed plane by plane and
    interpolated onto a Z_arrays mesh aligned with the grid:
    sum(linechg)*dz is the total charge q*w*np, and each plane sum is the
    plane's charge density times 1/(dx*dy) of its line charge."""
    rng = np.random.default_rng(3)
    nx, ny, nz = 9, 7, 12
    dx, dy, dz = 0.1, 0.2, 0.05
    x0, y0, z0 = -0.45, -0.7, 0.1
    n = 5000
    x = x0 + rng.uniform(0, nx * dx, n)
    y = y0 + rng.uniform(0, ny * dy, n)
    z = z0 + rng.uniform(0, nz * dz, n)
    q, w = 1.6e-19, 3.0e5
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O.setrho3d(rho, x, y, z, q, w, dx, dy, dz, x0, y0, z0)
    lc, planes = O.gtlchg3d(rho, dx, dy, dz, 0., z0, z0, 0., dz, nz)
    assert lc.sum() * dz == pytest.approx(q * w * n, rel=1e-12)
    np.testing.assert_allclose(planes, rho.sum(axis=(1, 2)), rtol=1e-13, atol=0)
    # iz*dzz/dz is iz up to rounding, so the weights are 1 and 0 to ~1e-16
    np.testing.assert_allclose(lc, planes * dx * dy, rtol=1e-14, atol=0)


@pytest.mark.parametrize("zbeam,nzzarr,dzz_over_dz,zgrid", [
    (0.013, 12, 1.0, 0.0), (-0.021, 27, 0.6, 0.0), (0.004, 7, 2.3, 0.031),
    (0.9, 12, 1.0, 0.0)])
def test_line_charge_interpolation_is_linear_in_the_beam_frame(zbeam, nzzarr, dzz_over_dz,
                                                               zgrid):
    """Point iz of the Z_arrays mesh sits at zzmin + iz*dzz in the beam
    frame, zbeam + zzmin + iz*dzz in the lab and (that - zgrid - zmmin)/dz
    in grid units; its line charge is the plane charges times dx*dy,
    linearly interpolated there: zero below the first plane and tapering to
    zero over the cell above the last (dw3d.F:305-313)."""
    rng = np.random.default_rng(17)
    nz, dx, dy, dz = 12, 0.1, 0.2, 0.05
    zmmin, zzmin = 0.1, 0.08
    planes = rng.uniform(0.5, 2.0, nz + 1)
    dzz = dzz_over_dz * dz
    lc = O.linechg_from_planes(planes, dx, dy, dz, zgrid, zmmin, zzmin, zbeam, dzz, nzzarr)
    zg = (zbeam + zzmin + np.arange(nzzarr + 1) * dzz - zgrid - zmmin) / dz
    want = np.interp(zg, np.arange(nz + 2), np.append(planes * dx * dy, 0.), left=0., right=0.)
    np.testing.assert_allclose(lc, want, rtol=0, atol=1e-14 * (planes * dx * dy).max())
    if zbeam == 0.9:
        assert not lc.any()


# ---------------------------------------------------------------------------
# particle boundaries, clearpart (top.F:3499-3800, 4730-4893) and getgrid3d
# (dtop.F:2818-2920)

def _ids_columns(alive):
    """Columns whose values name the particle (xp = index) with gaminv set
    from alive, uzp = index + 0.5."""
    n = len(alive)
    idx = np.arange(n, dtype=np.float64)
    cols = [idx.copy(), idx * 2., idx * 3., idx * 4., idx * 5., idx + 0.5,
            np.where(alive, 1.0, 0.0), idx * 7.]
    return cols


def test_clearpart_hand_example():
    """The worked example of clearpart's fill methods: live, lost, live,
    lost, live, live.  Method 1 fills hole 1 with the last live particle
    (5) and hole 3 with the one before it (4); method 2 keeps the order."""
    alive = np.array([1, 0, 1, 0, 1, 1], bool)
    c = _ids_columns(alive)
    assert O.clearpart(c, 1) == 4
    assert list(c[0][:4]) == [0., 5., 2., 4.]
    # moved-from slots: uzp and gaminv cleared
    assert c[5][5] == 0. and c[6][5] == 0. and c[5][4] == 0. and c[6][4] == 0.
    c = _ids_columns(alive)
    assert O.clearpart(c, 2) == 4
    assert list(c[0][:4]) == [0., 2., 4., 5.]
    c = _ids_columns(np.zeros(5, bool))
    assert O.clearpart(c, 1) == 0
    c = _ids_columns(np.ones(5, bool))
    assert O.clearpart(c, 1) == 5 and list(c[0]) == [0., 1., 2., 3., 4.]


@pytest.mark.parametrize("seed,frac", [(0, 0.0), (1, 0.1), (2, 0.5), (3, 0.93), (4, 1.0)])
def test_clearpart_fill_rules(seed, frac):
    """Method 1 characterised without its loop: the live set is kept, live
    particles below the new end stay where they are, and the k-th hole
    below the end receives the k-th live particle counted from the back.
    Method 2: the live particles in their original order.  Every column
    moves together."""
    rng = np.random.default_rng(seed)
    n = 997
    alive = rng.random(n) >= frac
    L = int(alive.sum())
    c = _ids_columns(alive)
    assert O.clearpart(c, 1) == L
    ids = c[0][:L].astype(int)
    assert sorted(ids) == list(np.nonzero(alive)[0])
    holes = [i for i in range(L) if not alive[i]]
    donors = [i for i in range(n - 1, L - 1, -1) if alive[i]]
    assert len(holes) == len(donors)
    for i in range(L):
        if alive[i]:
            assert ids[i] == i
    for h, d in zip(holes, donors):
        assert ids[h] == d
    for k, m in ((1, 2.), (2, 3.), (3, 4.), (4, 5.), (7, 7.)):
        np.testing.assert_array_equal(c[k][:L], ids * m)
    np.testing.assert_array_equal(c[5][:L], ids + 0.5)
    assert (c[6][:L] == 1.).all()
    c = _ids_columns(alive)
    assert O.clearpart(c, 2) == L
    np.testing.assert_array_equal(c[0][:L], np.nonzero(alive)[0].astype(float))


def test_particleboundaries_faces():
    """The absorbing tests: a particle exactly on a lower or upper face is
    lost (x <= xmin, x >= xmax), one a rounding step inside is kept; z is
    compared after subtracting zgrid; the upper extent is xmmin+nx*dx."""
    nx, ny, nz, dx, dy, dz = 4, 4, 8, 0.25, 0.25, 0.5
    x0, y0, z0, zg = -0.5, -0.5, 0.0, 0.75
    xs = [x0, np.nextafter(x0, 1), x0 + nx * dx, np.nextafter(x0 + nx * dx, -1), 0., 0., 0., 0.,
          0., np.nan]
    ys = [0., 0., 0., 0., y0, y0 + ny * dy, 0., 0., 0., 0.]
    zs = [1., 1., 1., 1., 1., 1., z0 + zg, nz * dz + zg, np.nextafter(nz * dz + zg, 0), 1.]
    x, y, z = (np.array(v, float) for v in (xs, ys, zs))
    g = np.ones(len(x))
    O.particleboundaries3d(x, y, z, g, dx, dy, dz, nx, ny, nz, x0, y0, z0, zg)
    assert list(g) == [0., 1., 0., 1., 0., 0., 0., 0., 1., 1.]


def test_getgrid3d_reproduces_linear_fields():
    """Trilinear interpolation is exact for a linear field (to rounding),
    gives node values on nodes exactly, and leaves f outside the grid
    (including the upper faces, gx < nx) untouched."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 7, 5, 9
    lo, hi = (-0.3, -0.2, 0.1), (0.4, 0.3, 1.0)
    xs = [np.linspace(l, h, n + 1) for l, h, n in zip(lo, hi, (nx, ny, nz))]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    grid = 1.5 + 2.0 * X - 3.0 * Y + 0.5 * Z
    n = 3000
    x = rng.uniform(lo[0] - 0.1, hi[0] + 0.1, n)
    y = rng.uniform(lo[1] - 0.1, hi[1] + 0.1, n)
    z = rng.uniform(lo[2] - 0.1, hi[2] + 0.1, n)
    f = np.full(n, -7.0)
    O.getgrid3d(x, y, z, f, grid, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    inside = ((x - lo[0]) * (nx / (hi[0] - lo[0])) >= 0) & \
        ((x - lo[0]) * (nx / (hi[0] - lo[0])) < nx) & \
        ((y - lo[1]) * (ny / (hi[1] - lo[1])) >= 0) & \
        ((y - lo[1]) * (ny / (hi[1] - lo[1])) < ny) & \
        ((z - lo[2]) * (nz / (hi[2] - lo[2])) >= 0) & \
        ((z - lo[2]) * (nz / (hi[2] - lo[2])) < nz)
    assert inside.sum() > n // 3 and (~inside).sum() > 0
    np.testing.assert_allclose(f[inside], (1.5 + 2.0 * x - 3.0 * y + 0.5 * z)[inside],
                               rtol=0, atol=1e-13)
    assert (f[~inside] == -7.0).all()
    # nodes (interior ones, chosen by index) get the node value exactly
    i, j, k = 3, 2, 4
    fx = np.array([xs[0][i]]), np.array([xs[1][j]]), np.array([xs[2][k]])
    g1 = np.zeros(1)
    O.getgrid3d(*fx, g1, grid, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    gx = (fx[0][0] - lo[0]) * (nx / (hi[0] - lo[0]))
    if gx == i:  # the scaled position lands on the node
        assert g1[0] == grid[k, j, i] or abs(g1[0] - grid[k, j, i]) < 1e-14
    # l4symtry reflects x and y about 0: f(-x, -y) = f(x, y)
    f1, f2 = np.zeros(200), np.zeros(200)
    xa, ya, za = rng.uniform(0, 0.3, 200), rng.uniform(0, 0.2, 200), rng.uniform(0.2, 0.9, 200)
    O.getgrid3d(xa, ya, za, f1, grid, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], l4symtry=True)
    O.getgrid3d(-xa, -ya, za, f2, grid, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                l4symtry=True)
    np.testing.assert_array_equal(f1, f2)
    # l2symtry reflects y only
    O.getgrid3d(xa, -ya, za, f2, grid, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                l2symtry=True)
    np.testing.assert_array_equal(f1, f2)


def test_periodic_z_boundary_wraps_by_periods():
    """pbound0 = periodic (top.F:3762-3778): z below the mesh moves up and z
    at or above it moves down by whole periods into [zmin+zg, zmax+zg); x
    and y still absorb; gaminv is untouched by the wrap."""
    nx, ny, nz, dx, dy, dz = 4, 4, 8, 0.25, 0.25, 0.5
    x0, y0, z0, zg = -0.5, -0.5, 0.0, 0.3
    L = nz * dz
    z = np.array([z0 + zg - 0.1, z0 + zg + L, z0 + zg + 2.5 * L, z0 + zg - 3.2 * L, 1.0, 1.0])
    x = np.array([0., 0., 0., 0., 0.5, 0.])
    y = np.zeros(6)
    g = np.ones(6)
    O.particleboundaries3d(x, y, z, g, dx, dy, dz, nx, ny, nz, x0, y0, z0, zg, zperiodic=True)
    assert list(g) == [1., 1., 1., 1., 0., 1.]
    # (a particle exactly one period up lands on zmin+zg up to the rounding
    # of z - L, as in the reference)
    assert ((z - zg >= z0 - 1e-12) & (z - zg < z0 + L)).all()
    np.testing.assert_allclose(z[:4] - zg, [L - 0.1, 0.0, 0.5 * L, 0.8 * L], rtol=0, atol=1e-12)
