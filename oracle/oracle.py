"""CPU FP64 oracle for the electrostatic PIC hot path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It wraps liboracle.so (oc_mgrid.c, oc_particles.c,
oc_conductors.c, oc_correctede.c), a loop-for-loop C restatement of the reference's Fortran
(pywarp90/f3d_mgrid.F, w3d.F, f3d_conductors.F), and restates the Python
geometry bookkeeping of scripts/generateconductors.py (Grid.getmeshnew,
installconductor) that feeds it.

Array conventions are the reference's: phi is Fortran-ordered (-1:nx+1)^3,
rho (0:nx)^3.  In numpy these are C-contiguous arrays of shape
(nz+3, ny+3, nx+3) and (nz+1, ny+1, nx+1), which is the same memory layout.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle.so")
LARGEPOS = 1.0e36
MAXLEVELS = 101


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        _lib = C.CDLL(_LIBPATH)
        _declare(_lib)
    return _lib


_i64 = C.c_int64
_d = C.c_double
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int64)


class OcConductors(C.Structure):
    _fields_ = [
        ("levels", _i64),
        ("levelnx", _i64 * MAXLEVELS), ("levelny", _i64 * MAXLEVELS),
        ("levelnz", _i64 * MAXLEVELS),
        ("levellx", _d * MAXLEVELS), ("levelly", _d * MAXLEVELS),
        ("levellz", _d * MAXLEVELS),
        ("int_n", _i64), ("int_nmax", _i64),
        ("int_indx", _pi), ("int_volt", _pd), ("int_numb", _pi),
        ("int_ilevel", _pi), ("int_istart", _i64 * (MAXLEVELS + 1)),
        ("sub_n", _i64 * 2), ("sub_nmax", _i64 * 2),
        ("sub_indx", _pi * 2), ("sub_dels", _pd * 2), ("sub_volt", _pd * 2),
        ("sub_numb", _pi * 2), ("sub_ilevel", _pi * 2),
        ("sub_prevphi", _pd * 2),
        ("sub_istart", (_i64 * (MAXLEVELS + 1)) * 2),
    ]


class OcIntercepts(C.Structure):
    _fields_ = [
        ("xmmin", _d), ("ymmin", _d), ("zmmin", _d),
        ("dx", _d), ("dy", _d), ("dz", _d),
        ("nx", _i64), ("ny", _i64), ("nz", _i64),
        ("ix", _i64), ("iy", _i64), ("iz", _i64), ("mglevel", _i64),
        ("nxicpt", _i64), ("nyicpt", _i64), ("nzicpt", _i64),
        ("xintercepts", _pd), ("yintercepts", _pd), ("zintercepts", _pd),
        ("xvoltages", _pd), ("yvoltages", _pd), ("zvoltages", _pd),
        ("xcondids", _pi), ("ycondids", _pi), ("zcondids", _pi),
    ]


def _declare(L):
    P = C.c_void_p
    L.oc_eps0.restype = _d
    L.oc_multigrid3dsolve.restype = _i64
    L.oc_multigrid3dsolve.argtypes = [
        _i64, _i64, _i64, _d, _d, _d, P, P, P, _d, _i64, P, _i64, _i64, P,
        _d, _i64, _i64, _i64, _i64, P]
    L.oc_sorhalfpass3d.argtypes = [_i64, _i64, _i64, _i64, _i64, P, P, _d, _d,
                                   _d, P, _d, _i64, _i64, _i64, P]
    L.oc_sorpass3d.argtypes = [_i64, _i64, _i64, _i64, P, P, _d, _d, _d, P,
                               _d, _i64, _i64, _i64, P]
    L.oc_residual3d.argtypes = [_i64, _i64, _i64, _d, _d, _d, P, P, P, _i64, P,
                                _d, _i64, _i64, _i64, _i64, P]
    L.oc_restrict3d.argtypes = [_i64, _i64, _i64, P, _i64, _i64, _i64, P, _d, P]
    L.oc_expand3d.argtypes = [_i64, _i64, _i64, P, _i64, _i64, _i64, P, P]
    L.oc_applyboundaryconditions3d.argtypes = [_i64, _i64, _i64, _i64, _i64,
                                               _i64, P, P, _i64, _i64]
    L.oc_getnextcoarselevel3d.argtypes = [_i64, _i64, _i64, _d, _d, _d, P, P, P,
                                          P, P, P]
    L.oc_getmglevels.argtypes = [_i64, _i64, _i64, _d, _d, _d, P]
    L.oc_checkconductors.argtypes = [_i64, _i64, _i64, _d, _d, _d, P]
    L.oc_setrho3d.argtypes = [P, _i64, P, P, P, _d, _d, _d, _i64, _i64, _i64,
                              _d, _d, _d, _d, _d, _d]
    L.oc_applyrhoboundaryconditions3d.argtypes = [_i64, _i64, _i64, P, P]
    L.oc_sete3d.argtypes = [P, _i64, P, P, P, _d, _d, _d, _d, _d, _d, _d, _i64,
                            _i64, _i64, P, P, P]
    L.oc_epush3d.argtypes = [_i64, P, P, P, P, P, P, _d, _d, _d]
    L.oc_xpush3d.argtypes = [_i64, P, P, P, P, P, P, P, _d]
    L.oc_getselfe3d.argtypes = [P, _i64, _i64, _i64, P, _d, _d, _d, _i64]
    L.oc_sete_from_e3d.argtypes = [P, _i64, P, P, P, _d, _d, _d, _d, _d, _d, _d, _i64,
                                   _i64, _i64, P, P, P]
    L.oc_getzmmnt.argtypes = [_i64, P, P, P, P, P, P, P, _i64, P]
    L.oc_particleboundaries3d.argtypes = [_i64, P, P, P, P, _d, _d, _d, _d, _d, _d, _i64,
                                          _i64, _i64, _d, _i64]
    L.oc_clearpart.restype = _i64
    L.oc_clearpart.argtypes = [_i64, _i64, P, _i64, _i64, _i64]
    L.oc_getgrid3d.argtypes = [_i64, P, P, P, P, _i64, _i64, _i64, P, _d, _d, _d, _d, _d,
                               _d, _i64, _i64]
    L.oc_setup_icgrid.argtypes = [_i64, _i64, _i64, _i64, P, _i64, P, _i64, P, P]
    L.oc_efield_subgrid.argtypes = [P, _i64, _i64, _i64, _d, _d, _d, P, _i64, P, P, P, P, P]
    L.oc_fixefield.argtypes = [P, _i64, _i64, _i64, _i64, P]
    L.oc_sete3d_conductor.argtypes = [P, P, P, P, P, P, P, _i64, P, P, P, _d, _d, _d, _d, _d,
                                      _d, _d, _i64, _i64, _i64, _i64, P, P, P]
    L.oc_plane_sums.argtypes = [P, _i64, _i64, _i64, _i64, _i64, P]
    L.oc_linechg_from_planes.argtypes = [P, _i64, _d, _d, _d, _d, _d, _d, _d, _d, _i64, P]
    L.oc_gtlchg3d.argtypes = [P, _i64, _i64, _i64, _d, _d, _d, _d, _d, _d, _d, _d, _i64,
                              _i64, _i64, P, P]
    L.oc_getese3d.restype = _d
    L.oc_getese3d.argtypes = [P, P, _i64, _i64, _i64, _d, _d, _d, _i64, _i64]
    L.oc_conductordelfromintercepts.restype = _i64
    L.oc_conductordelfromintercepts.argtypes = [P, P, _d, _d, _i64, P]
    L.oc_zcylinder_intercepts.argtypes = [_d, _d, _d, _d, _d, P, _d, _i64]
    L.oc_zsrfrv_intercepts.argtypes = [_i64, P, P, _d, _d, _d, P, _i64, P]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def eps0():
    return lib().oc_eps0()


# ---------------------------------------------------------------------------
# Conductors
# ---------------------------------------------------------------------------
class Conductors:
    """Holds the numpy storage behind an OcConductors struct."""

    def __init__(self):
        self.s = OcConductors()
        self.s.levels = 0
        for l in range(MAXLEVELS):
            self.s.levellx[l] = self.s.levelly[l] = self.s.levellz[l] = 1.0
        self.int_indx = np.zeros((0, 3), np.int64)
        self.int_volt = np.zeros(0)
        self.int_numb = np.zeros(0, np.int64)
        self.int_ilevel = np.zeros(0, np.int64)
        self.sub = [dict(indx=np.zeros((0, 3), np.int64), dels=np.zeros((0, 6)),
                         volt=np.zeros((0, 6)), numb=np.zeros((0, 6), np.int64),
                         ilevel=np.zeros(0, np.int64), prevphi=np.zeros(0))
                    for _ in range(2)]
        self._bind()

    def _bind(self):
        s = self.s
        s.int_nmax = len(self.int_volt)
        s.int_indx = self.int_indx.ctypes.data_as(_pi)
        s.int_volt = self.int_volt.ctypes.data_as(_pd)
        s.int_numb = self.int_numb.ctypes.data_as(_pi)
        s.int_ilevel = self.int_ilevel.ctypes.data_as(_pi)
        for p in range(2):
            d = self.sub[p]
            s.sub_nmax[p] = len(d["ilevel"])
            s.sub_indx[p] = d["indx"].ctypes.data_as(_pi)
            s.sub_dels[p] = d["dels"].ctypes.data_as(_pd)
            s.sub_volt[p] = d["volt"].ctypes.data_as(_pd)
            s.sub_numb[p] = d["numb"].ctypes.data_as(_pi)
            s.sub_ilevel[p] = d["ilevel"].ctypes.data_as(_pi)
            s.sub_prevphi[p] = d["prevphi"].ctypes.data_as(_pd)

    def grow(self, nc, ne, no):
        """Make room for nc, ne, no more interior/even/odd points."""
        n = self.s.int_n
        def g(a, extra):
            b = np.zeros((a.shape[0] + extra,) + a.shape[1:], a.dtype)
            b[:a.shape[0]] = a
            return b
        self.int_indx = g(self.int_indx[:n], nc)
        self.int_volt = g(self.int_volt[:n], nc)
        self.int_numb = g(self.int_numb[:n], nc)
        self.int_ilevel = g(self.int_ilevel[:n], nc)
        for p, extra in ((0, ne), (1, no)):
            m = self.s.sub_n[p]
            d = self.sub[p]
            for k in d:
                d[k] = g(d[k][:m], extra)
        self._bind()

    def ptr(self):
        return C.addressof(self.s)

    def lists(self):
        """Return copies of the sorted lists (after checkconductors)."""
        s = self.s
        out = dict(levels=s.levels,
                   levelnx=[s.levelnx[l] for l in range(s.levels)],
                   levelny=[s.levelny[l] for l in range(s.levels)],
                   levelnz=[s.levelnz[l] for l in range(s.levels)],
                   int_indx=self.int_indx[:s.int_n].copy(),
                   int_volt=self.int_volt[:s.int_n].copy(),
                   int_istart=[s.int_istart[l] for l in range(s.levels + 1)])
        for p in range(2):
            m = s.sub_n[p]
            d = self.sub[p]
            out["sub%d_indx" % p] = d["indx"][:m].copy()
            out["sub%d_dels" % p] = d["dels"][:m].copy()
            out["sub%d_volt" % p] = d["volt"][:m].copy()
            out["sub%d_istart" % p] = [s.sub_istart[p][l] for l in range(s.levels + 1)]
        return out


def _nint(x):
    # Fortran NINT: round half away from zero
    return int(np.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def getmglevels(nx, ny, nz, dx, dy, dz):
    c = Conductors()
    lib().oc_getmglevels(nx, ny, nz, dx, dy, dz, c.ptr())
    s = c.s
    return [dict(nx=s.levelnx[l], ny=s.levelny[l], nz=s.levelnz[l],
                 lx=s.levellx[l], ly=s.levelly[l], lz=s.levellz[l])
            for l in range(s.levels)]


class _Geom:
    """One z-aligned object: kind in ('zcylinder', 'zcylinderout', 'zsrfrv')."""

    def __init__(self, kind, voltage, condid, xcent=0., ycent=0., zcent=0.,
                 **kw):
        self.kind, self.voltage, self.condid = kind, voltage, condid
        self.xcent, self.ycent, self.zcent = xcent, ycent, zcent
        self.kw = kw
        if kind == "zcylinder":
            r, L = kw["radius"], kw["length"]
            self.extent = ([xcent - r, ycent - r, zcent - L / 2.],
                           [xcent + r, ycent + r, zcent + L / 2.])
        elif kind == "zcylinderout":
            L = kw["length"]
            self.extent = ([-LARGEPOS, -LARGEPOS, zcent - L / 2.],
                           [LARGEPOS, LARGEPOS, zcent + L / 2.])
        else:
            rmax = max(kw["rsrf"])
            self.extent = ([xcent - rmax, ycent - rmax, zcent + kw["zmin"]],
                           [xcent + rmax, ycent + rmax, zcent + kw["zmax"]])


def ZCylinder(radius, length, voltage=0., xcent=0., ycent=0., zcent=0., condid=1):
    """scripts/generateconductors.py:3305 (ZCylinder)."""
    return _Geom("zcylinder", voltage, condid, xcent, ycent, zcent,
                 radius=radius, length=length)


def ZCylinderOut(radius, length, voltage=0., xcent=0., ycent=0., zcent=0., condid=1):
    """scripts/generateconductors.py:3494 (ZCylinderOut)."""
    return _Geom("zcylinderout", voltage, condid, xcent, ycent, zcent,
                 radius=radius, length=length)


def ZAnnulus(rmin, rmax, length, voltage=0., xcent=0., ycent=0., zcent=0., condid=1):
    """scripts/generateconductors.py:5821-5855: a ZSrfrvIn whose z data is in
    the lab frame with zcent = 0, and the closed r-z outline of
    generateconductors.py:5012-5018."""
    zmin = -length / 2. + zcent
    zmax = +length / 2. + zcent
    rsrf = [rmin, rmax, rmax, rmin, rmin]
    zsrf = [zmin, zmin, zmax, zmax, zmin]
    return _Geom("zsrfrv", voltage, condid, xcent, ycent, 0.,
                 rsrf=rsrf, zsrf=zsrf, zmin=zmin, zmax=zmax)


def _getmeshnew(g, lev, extent):
    """Grid.getmeshnew, scripts/generateconductors.py:2592-2635 (serial,
    zbeam = 0)."""
    dx = g["dx"] * lev["lx"]
    dy = g["dy"] * lev["ly"]
    dz = g["dz"] * lev["lz"]
    nxl, nyl, nzl = lev["nx"], lev["ny"], lev["nz"]
    xmin = max(g["xmin"], g["xmmin"] + 0 * dx)
    xmax = min(g["xmax"], g["xmmin"] + (0 + nxl) * dx)
    ymin = max(g["ymin"], g["ymmin"] + 0 * dy)
    ymax = min(g["ymax"], g["ymmin"] + (0 + nyl) * dy)
    zmin = max(g["zmin"], g["zmmin"] + 0 * dz + 0.)
    zmax = min(g["zmax"], g["zmmin"] + (0 + nzl) * dz + 0.)
    mins, maxs = extent
    if mins[0] - dx > xmin:
        xmin = int((mins[0] - dx - xmin) / dx) * dx + xmin
        if xmin >= xmax: xmin = xmax - 2 * dx
    if mins[1] - dy > ymin:
        ymin = int((mins[1] - dy - ymin) / dy) * dy + ymin
        if ymin >= ymax: ymin = ymax - 2 * dy
    if mins[2] - dz > zmin:
        zmin = int((mins[2] - dz - zmin) / dz) * dz + zmin
        if zmin >= zmax: zmin = zmax - 2 * dz
    if maxs[0] + dx < xmax:
        xmax = int((maxs[0] + dx - xmin) / dx) * dx + xmin
        if xmax < maxs[0] + dx: xmax += dx
        if xmax <= xmin: xmax = xmin + 2 * dx
    if maxs[1] + dy < ymax:
        ymax = int((maxs[1] + dy - ymin) / dy) * dy + ymin
        if ymax < maxs[1] + dy: ymax += dy
        if ymax <= ymin: ymax = ymin + 2 * dy
    if maxs[2] + dz < zmax:
        zmax = int((maxs[2] + dz - zmin) / dz) * dz + zmin
        if zmax < maxs[2] + dz: zmax += dz
        if zmax <= zmin: zmax = zmin + 2 * dz
    nx = _nint((xmax - xmin) / dx)
    ny = _nint((ymax - ymin) / dy)
    nz = _nint((zmax - zmin) / dz)
    ix = _nint((xmin - (g["xmmin"] + 0 * dx)) / dx)
    iy = _nint((ymin - (g["ymmin"] + 0 * dy)) / dy)
    iz = _nint((zmin - (g["zmmin"] + 0 * dz + 0.)) / dz)
    return xmin, ymin, zmin, dx, dy, dz, nx, ny, nz, ix, iy, iz


def _checkoverlap(g, lev, extent):
    """Grid.checkoverlap, generateconductors.py:2637-2660 (serial)."""
    dx = g["dx"] * lev["lx"]
    dy = g["dy"] * lev["ly"]
    dz = g["dz"] * lev["lz"]
    mins, maxs = extent
    xmin = max(mins[0], g["xmin"]); ymin = max(mins[1], g["ymin"])
    zmin = max(mins[2], g["zmin"])
    xmax = min(maxs[0], g["xmax"]); ymax = min(maxs[1], g["ymax"])
    zmax = min(maxs[2], g["zmax"])
    if zmin - dz > zmax or xmin - dx > xmax or ymin - dy > ymax:
        return False
    return True


def _intercepts(geom, mesh, mglevel):
    xmin, ymin, zmin, dx, dy, dz, nx, ny, nz, ix, iy, iz = mesh
    if geom.kind == "zcylinder":
        nxi = nyi = nzi = 2
    elif geom.kind == "zcylinderout":
        nxi = nyi = 4
        nzi = 2
    else:
        nseg = len(geom.kw["rsrf"]) - 1
        nxi = nyi = nzi = 2 * nseg
    st = OcIntercepts()
    st.xmmin, st.ymmin, st.zmmin = xmin, ymin, zmin
    st.dx, st.dy, st.dz = dx, dy, dz
    st.nx, st.ny, st.nz = nx, ny, nz
    st.ix, st.iy, st.iz = ix, iy, iz
    st.mglevel = mglevel
    st.nxicpt, st.nyicpt, st.nzicpt = nxi, nyi, nzi
    arrs = dict(
        x=np.zeros((nz + 1, ny + 1, nxi)), y=np.zeros((nz + 1, nx + 1, nyi)),
        z=np.zeros((ny + 1, nx + 1, nzi)))
    for k in "xyz":
        arrs[k + "v"] = np.full(arrs[k].shape, float(geom.voltage))
        arrs[k + "c"] = np.full(arrs[k].shape, int(geom.condid), np.int64)
    st.xintercepts = arrs["x"].ctypes.data_as(_pd)
    st.yintercepts = arrs["y"].ctypes.data_as(_pd)
    st.zintercepts = arrs["z"].ctypes.data_as(_pd)
    st.xvoltages = arrs["xv"].ctypes.data_as(_pd)
    st.yvoltages = arrs["yv"].ctypes.data_as(_pd)
    st.zvoltages = arrs["zv"].ctypes.data_as(_pd)
    st.xcondids = arrs["xc"].ctypes.data_as(_pi)
    st.ycondids = arrs["yc"].ctypes.data_as(_pi)
    st.zcondids = arrs["zc"].ctypes.data_as(_pi)
    fuzz = 1.e-13
    if nx >= 0 and ny >= 0 and nz >= 0:
        if geom.kind in ("zcylinder", "zcylinderout"):
            lib().oc_zcylinder_intercepts(
                geom.kw["radius"], geom.kw["length"], geom.xcent, geom.ycent,
                geom.zcent, C.addressof(st), fuzz,
                1 if geom.kind == "zcylinderout" else 0)
        else:
            r = np.array(geom.kw["rsrf"], np.float64)
            z = np.array(geom.kw["zsrf"], np.float64)
            cnt = np.zeros(1, np.int64)
            lib().oc_zsrfrv_intercepts(len(r), _p(r), _p(z), geom.xcent,
                                       geom.ycent, geom.zcent, C.addressof(st),
                                       0, _p(cnt))
    return st, arrs


def install_conductors(geoms, nx, ny, nz, xmmin, xmmax, ymmin, ymmax,
                       zmmin, zmmax, dfill=2.0, conductors=None, crop=None):
    """installconductor + Grid.getdatanew + GridIntercepts.installlist
    (generateconductors.py:135-215, 2750-2858, 1851-1948), serial; crop =
    (xmin, xmax, ymin, ymax, zmin, zmax), entries None for the mesh's own
    (installconductor's region arguments, generateconductors.py:2421-2426)."""
    if conductors is None:
        conductors = Conductors()
    g = dict(xmmin=xmmin, xmmax=xmmax, ymmin=ymmin, ymmax=ymmax, zmmin=zmmin,
             zmmax=zmmax, dx=(xmmax - xmmin) / nx, dy=(ymmax - ymmin) / ny,
             dz=(zmmax - zmmin) / nz)
    c = crop if crop is not None else (None,) * 6
    for k, v, dflt in zip(("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"), c,
                          (xmmin, xmmax, ymmin, ymmax, zmmin, zmmax)):
        g[k] = dflt if v is None else v
    levels = getmglevels(nx, ny, nz, g["dx"], g["dy"], g["dz"])
    fuzz = 1.e-13
    for geom in geoms:
        if not _checkoverlap(g, levels[-1], geom.extent):
            continue
        for mglevel, lev in enumerate(levels):
            mesh = _getmeshnew(g, lev, geom.extent)
            st, arrs = _intercepts(geom, mesh, mglevel)
            cnt = np.zeros(3, np.int64)
            lib().oc_conductordelfromintercepts(C.addressof(st), conductors.ptr(),
                                                dfill, fuzz, 1, _p(cnt))
            conductors.grow(int(cnt[0]), int(cnt[1]), int(cnt[2]))
            lib().oc_conductordelfromintercepts(C.addressof(st), conductors.ptr(),
                                                dfill, fuzz, 0, _p(cnt))
    return conductors


def checkconductors(conductors, nx, ny, nz, dx, dy, dz):
    lib().oc_checkconductors(nx, ny, nz, dx, dy, dz, conductors.ptr())
    return conductors


# ---------------------------------------------------------------------------
# Solver and particle kernels
# ---------------------------------------------------------------------------
def multigrid3dsolve(phi, rho, dx, dy, dz, bounds=(0,) * 6, mgparam=1.2,
                     mgform=1, mgmaxiters=100, mgmaxlevels=101, mgtol=1e-6,
                     downpasses=1, uppasses=1, lcndbndy=True, icndbndy=2,
                     conductors=None):
    """multigrid3dsolve in place on phi (nz+3,ny+3,nx+3) and rho
    (nz+1,ny+1,nx+1); returns (mgiters, mgerror)."""
    nz, ny, nx = rho.shape[0] - 1, rho.shape[1] - 1, rho.shape[2] - 1
    assert phi.shape == (nz + 3, ny + 3, nx + 3)
    assert phi.dtype == np.float64 and rho.dtype == np.float64
    assert phi.flags.c_contiguous and rho.flags.c_contiguous
    if conductors is None:
        conductors = Conductors()
    b = np.array(bounds, np.int64)
    it = np.zeros(1, np.int64)
    err = np.zeros(1)
    lib().oc_multigrid3dsolve(nx, ny, nz, dx, dy, dz, _p(phi), _p(rho), _p(b),
                              mgparam, mgform, _p(it), mgmaxiters, mgmaxlevels,
                              _p(err), mgtol, downpasses, uppasses,
                              int(lcndbndy), icndbndy, conductors.ptr())
    return int(it[0]), float(err[0])


def setrho3d(rho, xp, yp, zp, q, wght, dx, dy, dz, xmmin, ymmin, zmmin,
             zgrid=0.0):
    nz, ny, nx = rho.shape[0] - 1, rho.shape[1] - 1, rho.shape[2] - 1
    xp, yp, zp = _f64(xp), _f64(yp), _f64(zp)
    # The restated loop has no bounds check (neither has the reference);
    # refuse particles whose cell is outside the mesh.
    for v, lo, d, n in ((xp, xmmin, dx, nx), (yp, ymmin, dy, ny), (zp - zgrid, zmmin, dz, nz)):
        s = (v - lo) * (1. / d)
        if len(s) and not (s.min() >= 0. and s.max() < n):
            raise ValueError("setrho3d: particles outside the mesh")
    lib().oc_setrho3d(_p(rho), len(xp), _p(xp), _p(yp), _p(zp), zgrid, q, wght,
                      nx, ny, nz, dx, dy, dz, xmmin, ymmin, zmmin)


def applyrhoboundaryconditions3d(rho, bounds):
    nz, ny, nx = rho.shape[0] - 1, rho.shape[1] - 1, rho.shape[2] - 1
    b = np.array(bounds, np.int64)
    lib().oc_applyrhoboundaryconditions3d(nx, ny, nz, _p(rho), _p(b))


def sete3d(phi, xp, yp, zp, ex, ey, ez, dx, dy, dz, xmmin, ymmin, zmmin,
           zgrid=0.0):
    nz, ny, nx = phi.shape[0] - 3, phi.shape[1] - 3, phi.shape[2] - 3
    xp, yp, zp = _f64(xp), _f64(yp), _f64(zp)
    for v, lo, d, n in ((xp, xmmin, dx, nx), (yp, ymmin, dy, ny), (zp - zgrid, zmmin, dz, nz)):
        s = (v - lo) * (1. / d)
        if len(s) and not (s.min() >= 0. and s.max() < n):
            raise ValueError("sete3d: particles outside the mesh")
    lib().oc_sete3d(_p(phi), len(xp), _p(xp), _p(yp), _p(zp), zgrid, xmmin,
                    ymmin, zmmin, dx, dy, dz, nx, ny, nz, _p(ex), _p(ey), _p(ez))


def epush3d(uxp, uyp, uzp, ex, ey, ez, q, m, ddt):
    lib().oc_epush3d(len(uxp), _p(uxp), _p(uyp), _p(uzp), _p(ex), _p(ey),
                     _p(ez), q, m, ddt)


def xpush3d(xp, yp, zp, uxp, uyp, uzp, gaminv, dtp):
    lib().oc_xpush3d(len(xp), _p(xp), _p(yp), _p(zp), _p(uxp), _p(uyp),
                     _p(uzp), _p(gaminv), dtp)


def sete_from_e3d(selfe, xp, yp, zp, ex, ey, ez, dx, dy, dz, xmmin, ymmin, zmmin,
                  zgrid=0.0):
    """sete3d with efetch=3: trilinear interpolation of the grid field selfe
    (nz+1, ny+1, nx+1, 3), added to ex, ey, ez (w3d.F:4362-4460)."""
    nz, ny, nx = selfe.shape[0] - 1, selfe.shape[1] - 1, selfe.shape[2] - 1
    xp, yp, zp = _f64(xp), _f64(yp), _f64(zp)
    lib().oc_sete_from_e3d(_p(np.ascontiguousarray(selfe)), len(xp), _p(xp), _p(yp), _p(zp),
                           zgrid, xmmin, ymmin, zmmin, dx, dy, dz, nx, ny, nz, _p(ex), _p(ey),
                           _p(ez))


def getselfe3d(phi, dx, dy, dz, selfe=None, lzero=True):
    nz, ny, nx = phi.shape[0] - 3, phi.shape[1] - 3, phi.shape[2] - 3
    if selfe is None:
        selfe = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    lib().oc_getselfe3d(_p(phi), nx, ny, nz, _p(selfe), dx, dy, dz, int(lzero))
    return selfe


def getzmmnt(xp, yp, zp, uxp, uyp, uzp, gaminv, usegaminv=False):
    """The 33 whole-beam moment sums zmmnts0 of getzmmnt (top.F:1059-1136)."""
    mm = np.zeros(33)
    a = [_f64(v) for v in (xp, yp, zp, uxp, uyp, uzp, gaminv)]
    lib().oc_getzmmnt(len(a[0]), *[_p(v) for v in a], int(usegaminv), _p(mm))
    return mm


def particleboundaries3d(xp, yp, zp, gaminv, dx, dy, dz, nx, ny, nz, xmmin, ymmin, zmmin,
                         zgrid=0.0, zperiodic=False):
    """Flags particles outside the absorbing mesh boundaries with gaminv = 0
    (particleboundaries3d, w3d.F:11292-11450); with zperiodic z wraps by
    whole periods instead of absorbing."""
    lib().oc_particleboundaries3d(len(xp), _p(xp), _p(yp), _p(zp), _p(gaminv), xmmin, ymmin,
                                  zmmin, dx, dy, dz, nx, ny, nz, zgrid, int(zperiodic))


def clearpart(cols, fillmethod=1):
    """clearpart (top.F:4730-4893) in place on the columns
    [xp, yp, zp, uxp, uyp, uzp, gaminv, ...] (float64 arrays of one length);
    returns the number of live particles now at the front."""
    cols = list(cols)
    for c in cols:
        assert c.dtype == np.float64 and c.flags.c_contiguous and len(c) == len(cols[0])
    arr = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    return int(lib().oc_clearpart(len(cols[0]), len(cols), arr, 5, 6, fillmethod))


def getgrid3d(x, y, z, f, grid, xmin, xmax, ymin, ymax, zmin, zmax, l2symtry=False,
              l4symtry=False):
    """getgrid3d (dtop.F:2818-2920): f at the positions inside the grid
    (shape (nz+1, ny+1, nx+1)) by linear interpolation; f unchanged elsewhere."""
    nz, ny, nx = grid.shape[0] - 1, grid.shape[1] - 1, grid.shape[2] - 1
    x, y, z = _f64(x), _f64(y), _f64(z)
    g = np.ascontiguousarray(grid, dtype=np.float64)
    lib().oc_getgrid3d(len(x), _p(x), _p(y), _p(z), _p(f), nx, ny, nz, _p(g), xmin, xmax,
                       ymin, ymax, zmin, zmax, int(l2symtry), int(l4symtry))


def plane_sums(rho, l2symtry=False, l4symtry=False):
    """The charge of each z plane of rho (nz+1, ny+1, nx+1), gtlchg3dfromrho's
    scrtch (dw3d.F:254-279)."""
    nz, ny, nx = (d - 1 for d in rho.shape)
    out = np.zeros(nz + 1)
    lib().oc_plane_sums(_p(_f64(rho)), nx, ny, nz, int(l2symtry), int(l4symtry), _p(out))
    return out


def linechg_from_planes(planes, dx, dy, dz, zgrid, zmmin, zzmin, zbeam, dzz, nzzarr):
    """dw3d.F:282-316: plane sums -> linechg(0:nzzarr)."""
    planes = _f64(planes)
    out = np.zeros(nzzarr + 1)
    lib().oc_linechg_from_planes(_p(planes), len(planes) - 1, dx, dy, dz, zgrid, zmmin,
                                 zzmin, zbeam, dzz, nzzarr, _p(out))
    return out


def gtlchg3d(rho, dx, dy, dz, zgrid, zmmin, zzmin, zbeam, dzz, nzzarr, l2symtry=False,
             l4symtry=False):
    """gtlchg3dfromrho (dw3d.F:229-320): returns (linechg, plane sums)."""
    rho = _f64(rho)
    nz, ny, nx = (d - 1 for d in rho.shape)
    scr = np.zeros(nz + 1)
    lc = np.zeros(nzzarr + 1)
    lib().oc_gtlchg3d(_p(rho), nx, ny, nz, dx, dy, dz, zgrid, zmmin, zzmin, zbeam, dzz,
                      nzzarr, int(l2symtry), int(l4symtry), _p(scr), _p(lc))
    return lc, scr


def getese3d(rho, phi, dx, dy, dz, l4symtry=False, l2symtry=False):
    """getese3dfromrhophi (dw3d.F:708-768): electrostatic energy of rho
    (nz+1, ny+1, nx+1) in phi (nz+3, ny+3, nx+3)."""
    nz, ny, nx = (d - 1 for d in rho.shape)
    assert phi.shape == (nz + 3, ny + 3, nx + 3)
    return float(lib().oc_getese3d(_p(_f64(rho)), _p(_f64(phi)), nx, ny, nz, dx, dy, dz,
                                   int(l4symtry), int(l2symtry)))


def sorhalfpass3d(parity, phi, rho, dx, dy, dz, bounds, mgparam=1.2, mgform=1,
                  mglevel=0, conductors=None, lcndbndy=True, icndbndy=2):
    nz, ny, nx = rho.shape[0] - 1, rho.shape[1] - 1, rho.shape[2] - 1
    b = np.array(bounds, np.int64)
    cp = conductors.ptr() if conductors is not None else None
    lib().oc_sorhalfpass3d(parity, mglevel, nx, ny, nz, _p(phi), _p(rho),
                           1. / (dx * dx), 1. / (dy * dy), 1. / (dz * dz), _p(b),
                           mgparam, mgform, int(lcndbndy), icndbndy, cp)


def residual3d(phi, rho, dx, dy, dz, bounds, mgparam=1.2, mgform=1, mglevel=0,
               conductors=None, lcndbndy=True, icndbndy=2, mgform2init=False):
    nz, ny, nx = rho.shape[0] - 1, rho.shape[1] - 1, rho.shape[2] - 1
    res = np.zeros((nz + 3, ny + 3, nx + 3))
    b = np.array(bounds, np.int64)
    cp = conductors.ptr() if conductors is not None else None
    lib().oc_residual3d(nx, ny, nz, 1. / (dx * dx), 1. / (dy * dy), 1. / (dz * dz),
                        _p(phi), _p(rho), _p(res), mglevel, _p(b), mgparam, mgform,
                        int(mgform2init), int(lcndbndy), icndbndy, cp)
    return res


def getnextcoarselevel3d(nx, ny, nz, dx, dy, dz):
    n = np.zeros(3, np.int64)
    d = np.zeros(3)
    lib().oc_getnextcoarselevel3d(nx, ny, nz, dx, dy, dz, _p(n[0:]), _p(n[1:]),
                                  _p(n[2:]), _p(d[0:]), _p(d[1:]), _p(d[2:]))
    return int(n[0]), int(n[1]), int(n[2]), float(d[0]), float(d[1]), float(d[2])


def restrict3d(res, nxc, nyc, nzc, ff, bounds):
    nz, ny, nx = res.shape[0] - 3, res.shape[1] - 3, res.shape[2] - 3
    rc = np.zeros((nzc + 1, nyc + 1, nxc + 1))
    b = np.array(bounds, np.int64)
    lib().oc_restrict3d(nx, ny, nz, _p(res), nxc, nyc, nzc, _p(rc), ff, _p(b))
    return rc


def expand3d(phi, phicoarse, bounds):
    nz, ny, nx = phi.shape[0] - 3, phi.shape[1] - 3, phi.shape[2] - 3
    nzc, nyc, nxc = (phicoarse.shape[0] - 3, phicoarse.shape[1] - 3,
                     phicoarse.shape[2] - 3)
    b = np.array(bounds, np.int64)
    lib().oc_expand3d(nx, ny, nz, _p(phi), nxc, nyc, nzc, _p(phicoarse), _p(b))


def applyboundaryconditions3d(u, bounds, lwithdirichlet=False,
                              lzerodirichlet=False):
    nz, ny, nx = u.shape[0] - 3, u.shape[1] - 3, u.shape[2] - 3
    b = np.array(bounds, np.int64)
    lib().oc_applyboundaryconditions3d(nx, ny, nz, 1, 1, 1, _p(u), _p(b),
                                       int(lwithdirichlet), int(lzerodirichlet))


# ---------------------------------------------------------------------------
# Conductor-corrected E field (f3d.lcorrectede)
# ---------------------------------------------------------------------------
OC_INTERIOR = 2147483647


class CondField:
    """setupconductorfielddata (f3d_mgrid.F:4581-4633) for the finest level:
    the node classification icgrid (nz+1, ny+1, nx+1) and the per-record
    E arrays of both parity lists, from a lists() dict of Conductors."""

    def __init__(self, lists, nx, ny, nz):
        self.nx, self.ny, self.nz = nx, ny, nz

        def lev0(key, start):
            a = np.ascontiguousarray(lists[key], dtype=np.int64 if "indx" in key else np.float64)
            st = lists[start]
            return a[st[0]:st[1]] if len(st) > 1 else a[:0]
        self.int_indx = lev0("int_indx", "int_istart").reshape(-1, 3).copy()
        self.sub = []
        for p in range(2):
            self.sub.append(dict(indx=lev0("sub%d_indx" % p, "sub%d_istart" % p).reshape(-1, 3).copy(),
                                 dels=lev0("sub%d_dels" % p, "sub%d_istart" % p).reshape(-1, 6).copy(),
                                 volt=lev0("sub%d_volt" % p, "sub%d_istart" % p).reshape(-1, 6).copy()))
        # the reference skips the subgrid E when either whole list (all
        # levels) is empty (f3d_mgrid.F:4659-4661)
        self.skip = len(lists["sub0_indx"]) == 0 or len(lists["sub1_indx"]) == 0
        self.icgrid = np.zeros((nz + 1, ny + 1, nx + 1), np.int32)
        e, o = self.sub
        lib().oc_setup_icgrid(nx, ny, nz, len(self.int_indx), _p(self.int_indx), len(e["indx"]),
                              _p(e["indx"]), len(o["indx"]), _p(o["indx"]), _p(self.icgrid))
        for d in self.sub:
            n = len(d["indx"])
            d["efield0"] = np.zeros((n, 3))
            d["efieldd"] = np.zeros((n, 6))

    def efield(self, phi, dx, dy, dz):
        """getefieldatconductorsubgrid (f3d_mgrid.F:4635-4735) for the
        potential phi (nz+3, ny+3, nx+3)."""
        if self.skip:
            return
        assert phi.shape == (self.nz + 3, self.ny + 3, self.nx + 3) and phi.flags.c_contiguous
        for d in self.sub:
            lib().oc_efield_subgrid(_p(phi), self.nx, self.ny, self.nz, dx, dy, dz,
                                    _p(self.icgrid), len(d["indx"]), _p(d["indx"]),
                                    _p(d["dels"]), _p(d["volt"]), _p(d["efield0"]),
                                    _p(d["efieldd"]))

    def fixefield(self, selfe):
        """fixefieldatconductorpoints (f3d_mgrid.F:4737-4772) in place on
        selfe (nz+1, ny+1, nx+1, 3)."""
        if len(self.int_indx) == 0:
            return
        assert selfe.flags.c_contiguous
        lib().oc_fixefield(_p(selfe), self.nx, self.ny, self.nz, len(self.int_indx),
                           _p(self.int_indx))


def sete3d_conductor(cf, phi, selfe, xp, yp, zp, ex, ey, ez, dx, dy, dz, xmmin, ymmin, zmmin,
                     efetch=1, zgrid=0.0):
    """sete3dwithconductor (f3d_mgrid.F:4774-5010, 5164-5233) for efetch 1
    (from phi) or 3 (from selfe): E added to ex, ey, ez."""
    if efetch not in (1, 3):
        raise ValueError("efetch must be 1 or 3")
    xp, yp, zp = _f64(xp), _f64(yp), _f64(zp)
    e, o = cf.sub
    z = np.zeros(1)
    sp = _p(np.ascontiguousarray(selfe)) if selfe is not None else _p(z)
    if efetch == 3 and selfe is None:
        raise ValueError("efetch 3 needs selfe")
    lib().oc_sete3d_conductor(_p(cf.icgrid), _p(e["efield0"]), _p(e["efieldd"]),
                              _p(o["efield0"]), _p(o["efieldd"]), _p(phi), sp, len(xp), _p(xp),
                              _p(yp), _p(zp), zgrid, xmmin, ymmin, zmmin, dx, dy, dz, cf.nx,
                              cf.ny, cf.nz, efetch, _p(ex), _p(ey), _p(ez))
