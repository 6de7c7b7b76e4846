"""Shared builders for the tests: manufactured solutions, seeded loads,
and the configuration geometries."""
import numpy as np

from oracle import oracle as O


def manufactured(nx, ny, nz, bounds, lo=(-1., -1., -1.), hi=(1., 1., 1.)):
    """The reference's multigrid test problem (warp_test/MutltiGrid3D_test.py
    with fieldsolvertesttools.gettestrho): phi = sin(pi x+.5) sin(2 pi y+2.5)
    sin(pi z+1.111) with guard cells edited per boundary type, and rho =
    -eps0 * (7-point Laplacian of phi).  The discrete solution is phi.
    Returns (phi_exact, rho, phi_initial, dx, dy, dz)."""
    dx, dy, dz = [(h - l) / n for l, h, n in zip(lo, hi, (nx, ny, nz))]
    x = lo[0] - dx + dx * np.arange(nx + 3)
    y = lo[1] - dy + dy * np.arange(ny + 3)
    z = lo[2] - dz + dz * np.arange(nz + 3)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    phi = np.sin(np.pi * X + 0.5) * np.sin(2 * np.pi * Y + 2.5) * np.sin(np.pi * Z + 1.111)
    c = slice(1, -1)
    b = bounds
    if b[0] == 1: phi[c, c, 0] = phi[c, c, 2]
    if b[1] == 1: phi[c, c, -1] = phi[c, c, -3]
    if b[2] == 1: phi[c, 0, c] = phi[c, 2, c]
    if b[3] == 1: phi[c, -1, c] = phi[c, -3, c]
    if b[4] == 1: phi[0, c, c] = phi[2, c, c]
    if b[5] == 1: phi[-1, c, c] = phi[-3, c, c]
    if b[0] == 0: phi[c, c, 0] = 2 * phi[c, c, 1] - phi[c, c, 2]
    if b[1] == 0: phi[c, c, -1] = 2 * phi[c, c, -2] - phi[c, c, -3]
    if b[2] == 0: phi[c, 0, c] = 2 * phi[c, 1, c] - phi[c, 2, c]
    if b[3] == 0: phi[c, -1, c] = 2 * phi[c, -2, c] - phi[c, -3, c]
    if b[4] == 0: phi[0, c, c] = 2 * phi[1, c, c] - phi[2, c, c]
    if b[5] == 0: phi[-1, c, c] = 2 * phi[-2, c, c] - phi[-3, c, c]
    e0 = O.eps0()
    rho = -e0 * ((phi[c, c, :-2] - 2 * phi[c, c, c] + phi[c, c, 2:]) / dx ** 2 +
                 (phi[c, :-2, c] - 2 * phi[c, c, c] + phi[c, 2:, c]) / dy ** 2 +
                 (phi[:-2, c, c] - 2 * phi[c, c, c] + phi[2:, c, c]) / dz ** 2)
    phi0 = np.zeros_like(phi)
    if b[0] == 0: phi0[:, :, 1] = phi[:, :, 1]
    if b[1] == 0: phi0[:, :, -2] = phi[:, :, -2]
    if b[2] == 0: phi0[:, 1, :] = phi[:, 1, :]
    if b[3] == 0: phi0[:, -2, :] = phi[:, -2, :]
    if b[4] == 0: phi0[1, :, :] = phi[1, :, :]
    if b[5] == 0: phi0[-2, :, :] = phi[-2, :, :]
    return phi, np.ascontiguousarray(rho), phi0, dx, dy, dz


def clearcorners(p):
    """fieldsolvertesttools.clearcorners: the edge guard cells are not
    defined by the solver."""
    p[:, 0, 0] = 0; p[:, 0, -1] = 0; p[:, -1, 0] = 0; p[:, -1, -1] = 0
    p[0, :, 0] = 0; p[0, :, -1] = 0; p[-1, :, 0] = 0; p[-1, :, -1] = 0
    p[0, 0, :] = 0; p[0, -1, :] = 0; p[-1, 0, :] = 0; p[-1, -1, :] = 0


def all_bounds():
    """Every boundary combination of MutltiGrid3D_test.py:57-83."""
    seen = []
    for zp in [0, 1, 2]:
        for zm in [0, 1, 2]:
            if zp == 2: zm = 2
            if zm == 2: zp = 2
            for yp in [0, 1, 2]:
                for ym in [0, 1, 2]:
                    if yp == 2: ym = 2
                    if ym == 2: yp = 2
                    for xp in [0, 1, 2]:
                        for xm in [0, 1, 2]:
                            if xp == 2: xm = 2
                            if xm == 2: xp = 2
                            b = (xm, xp, ym, yp, zm, zp)
                            if b not in seen:
                                seen.append(b)
    return seen


# Configuration geometry (domain x,y in [-2,2] cm, z in [0,8] cm).
DOMAIN = (-0.02, 0.02, -0.02, 0.02, 0.0, 0.08)
PIPE_R = 0.018
PIPE_LEN = 0.10       # centred at z = 4 cm: covers the mesh with margin
ANN_RIN, ANN_ROUT = 0.012, 0.018
ANN_Z0, ANN_Z1 = 0.03, 0.035
ANN_V = 1.0e4


def pipe_and_annulus(mod):
    """C1's conductors with the classes of module mod (oracle or warp_amd)."""
    return [mod.ZCylinderOut(PIPE_R, PIPE_LEN, 0., zcent=0.04, condid=1),
            mod.ZAnnulus(ANN_RIN, ANN_ROUT, ANN_Z1 - ANN_Z0, ANN_V,
                         zcent=0.5 * (ANN_Z0 + ANN_Z1), condid=2)]


def slab_lists(n3, axis, side, ic, frac, V, dfill=2.0):
    """Conductor lists (level 0 only) of a slab at voltage V normal to
    `axis` on a grid of n3 = (nx, ny, nz) cells.  side = -1: the slab fills
    s <= (ic - frac) ds, covering the nodes of index < ic; side = +1: it
    fills s >= (ic + frac) ds, covering the nodes of index > ic.  The nodes
    of index ic are subgrid points whose neighbour across the surface is
    covered (dels = frac there, dfill in the five open directions)."""
    nx, ny, nz = n3
    nodes = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                                 indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    s = nodes[:, axis]
    cov = s < ic if side < 0 else s > ic
    out = dict(levels=1, int_indx=np.ascontiguousarray(nodes[cov]).astype(np.int64),
               int_volt=np.full(int(cov.sum()), float(V)),
               int_istart=np.array([0, int(cov.sum())], np.int64))
    k = 2 * axis + (0 if side < 0 else 1)
    for p in range(2):
        pts = nodes[(s == ic) & (nodes.sum(1) % 2 == p)].astype(np.int64)
        dels = np.full((len(pts), 6), dfill)
        dels[:, k] = frac
        volt = np.zeros((len(pts), 6))
        volt[:, k] = V
        out["sub%d_indx" % p] = np.ascontiguousarray(pts)
        out["sub%d_dels" % p] = dels
        out["sub%d_volt" % p] = volt
        out["sub%d_istart" % p] = np.array([0, len(pts)], np.int64)
    return out


def neumann_lists(n3, ic, kind, frac, vfun, levels=1):
    """Hand-made subgrid lists of a sheet normal to x through the node plane
    ix = ic (level 0; the coarser levels, up to `levels`, hold none).
    Every node of the plane is a subgrid point whose x directions are, by
    kind: "m" Neumann below (dels[mx] = -frac), "p" Neumann above, "mp"
    Neumann on both sides, "dm" Dirichlet below (dels[mx] = frac at the
    voltage vfun(ix, iy, iz)) and Neumann above; the other four directions
    are bulk (dels = 2).  No interior points: a Neumann conductor installs
    none (Delta.setparity with dfill = 0, generateconductors.py:1623-1626).
    The dels signs select the branches of handlesubgrid2
    (f3d_mgrid.F:2661-2762)."""
    nx, ny, nz = n3
    nodes = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                                 indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    mx, px = {"m": (-frac, 2.), "p": (2., -frac), "mp": (-frac, -frac),
              "dm": (frac, -frac)}[kind]
    out = dict(levels=levels, int_indx=np.zeros((0, 3), np.int64), int_volt=np.zeros(0),
               int_istart=np.zeros(levels + 1, np.int64))
    for p in range(2):
        pts = nodes[(nodes[:, 0] == ic) & (nodes.sum(1) % 2 == p)].astype(np.int64)
        dels = np.full((len(pts), 6), 2.)
        dels[:, 0], dels[:, 1] = mx, px
        volt = np.zeros((len(pts), 6))
        if kind == "dm":
            volt[:, 0] = [vfun(*q) for q in pts]
        out["sub%d_indx" % p] = np.ascontiguousarray(pts)
        out["sub%d_dels" % p] = dels
        out["sub%d_volt" % p] = volt
        out["sub%d_istart" % p] = np.array([0] + [len(pts)] * levels, np.int64)
    return out


def oracle_conductors(d):
    """The oracle's Conductors holding the level-0 lists of a lists() dict
    (condid 1); multigrid3dsolve sorts them per level (checkconductors)."""
    from oracle import oracle as O
    co = O.Conductors()
    n = [len(d["int_indx"]), len(d["sub0_indx"]), len(d["sub1_indx"])]
    co.grow(*n)
    co.int_indx[:] = d["int_indx"]
    co.int_volt[:] = d["int_volt"]
    co.int_numb[:] = 1
    co.int_ilevel[:] = 0
    co.s.int_n = n[0]
    for p in range(2):
        q = co.sub[p]
        q["indx"][:] = d["sub%d_indx" % p]
        q["dels"][:] = d["sub%d_dels" % p]
        q["volt"][:] = d["sub%d_volt" % p]
        q["numb"][:] = 1
        q["ilevel"][:] = 0
        q["prevphi"][:] = 0.
        co.s.sub_n[p] = n[1 + p]
    return co


def linear_z_case(n3, d3, g, h, seed):
    """phi = g + h z on every node and guard (the boundary values of six
    Dirichlet faces), with the interior nodes perturbed by uniform noise in
    [-1, 1); returns (initial phi, exact phi), both (nz+3, ny+3, nx+3)."""
    nx, ny, nz = n3
    z = (np.arange(-1, nz + 2) * d3[2])[:, None, None]
    exact = np.ascontiguousarray(np.broadcast_to(g + h * z, (nz + 3, ny + 3, nx + 3)))
    phi = exact.copy()
    rng = np.random.default_rng(seed)
    phi[2:nz + 1, 2:ny + 1, 2:nx + 1] += rng.uniform(-1., 1., (nz - 1, ny - 1, nx - 1))
    return phi, exact


def slab_surface(axis, side, ic, frac, d3):
    """Position of the slab's surface along its axis."""
    return (ic + side * frac) * d3[axis]


def slab_f(t, coef):
    g, h, q = coef
    return g + h * t + q * t * t


def slab_phi(n3, d3, axis, side, ic, frac, V, coef):
    """phi (nz+3, ny+3, nx+3) of the slab: V on the covered nodes, and
    V + (s - s_c) f(t) elsewhere, with s along the axis, s_c the surface,
    t the next axis' coordinate and f(t) = g + h t + q t^2."""
    nx, ny, nz = n3
    c = [np.arange(-1, n + 2) for n in n3]
    I = np.meshgrid(*c, indexing="ij")          # (nx+3, ny+3, nz+3)
    si = I[axis]
    s = si * d3[axis]
    ta = (axis + 1) % 3
    t = I[ta] * d3[ta]
    sc = slab_surface(axis, side, ic, frac, d3)
    cov = si < ic if side < 0 else si > ic
    phi = np.where(cov, float(V), V + (s - sc) * slab_f(t, coef))
    return np.ascontiguousarray(phi.transpose(2, 1, 0))


def lists_struct(d):
    """A WmConductorLists over the arrays of a lists() dict (kept alive on
    the returned struct)."""
    import ctypes as C
    from warp_amd import _native as N
    v = N.WmConductorLists()
    keep = []

    def P(a, t):
        a = np.ascontiguousarray(a, dtype=np.int64 if t is C.c_int64 else np.float64).reshape(-1)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(t))
    v.levels = int(d["levels"])
    v.int_n = len(d["int_indx"])
    v.int_indx = P(d["int_indx"], C.c_int64)
    v.int_volt = P(d["int_volt"], C.c_double)
    v.int_istart = P(d["int_istart"], C.c_int64)
    for p in range(2):
        v.sub_n[p] = len(d["sub%d_indx" % p])
        v.sub_indx[p] = P(d["sub%d_indx" % p], C.c_int64)
        v.sub_dels[p] = P(d["sub%d_dels" % p], C.c_double)
        v.sub_volt[p] = P(d["sub%d_volt" % p], C.c_double)
        v.sub_istart[p] = P(d["sub%d_istart" % p], C.c_int64)
    v._keep = keep
    return v
