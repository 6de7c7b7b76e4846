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
