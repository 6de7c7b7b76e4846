"""Conductor geometry and installation, mirroring scripts/generateconductors.py.

Objects (ZCylinder, ZCylinderOut, ZAnnulus, ZSrfrvIn with straight
segments, and sums of them) produce grid-line intercepts on every multigrid
level (Grid.getdatanew, generateconductors.py:2750-2808); libwarpmi turns
them into interior points and subgrid distances
(conductordelfromintercepts, f3d_conductors.F:12-728) and cleans and sorts
them per level (checkconductors, f3d_mgrid.F:3964-4326).  The result is a
:class:`ConductorData`, the counterpart of the reference's ConductorType,
which the solver uploads to the GPU.
"""
import ctypes as C

import numpy as np

from . import _native as N

LARGEPOS = 1.0e36  # pywarp90/top.h:3


class ConductorExtent:
    """generateconductors.py ConductorExtent: axis-aligned bounds."""

    def __init__(self, mins, maxs):
        self.mins = list(mins)
        self.maxs = list(maxs)


class Assembly:
    """Base of all conductors (generateconductors.py:230-).

    voltage, xcent, ycent, zcent and condid as in the reference; condid
    'next' takes the next unused integer id.  neumann is kept as the
    reference keeps it: its intercept generation, the default one
    (_lwithnewconductorgeneration = True, generateconductors.py:124, 200-205),
    hands the flag to GridIntercepts (generateconductors.py:333-340,
    1815-1821), which stores nothing of it, and conductordelfromintercepts
    takes no such argument, so the installed data do not depend on it.
    Only the older distance-based generation (Delta.normalize and
    setparity, generateconductors.py:1551-1627) reads it; that generation
    is not provided.
    """
    nextcondid = 1

    def __init__(self, voltage=0., xcent=0., ycent=0., zcent=0., condid='next', neumann=0):
        self.neumann = neumann
        self.voltage = voltage
        self.xcent = xcent
        self.ycent = ycent
        self.zcent = zcent
        if condid == 'next':
            self.condid = Assembly.nextcondid
            Assembly.nextcondid += 1
        else:
            self.condid = condid
        self.extent = None

    def createextent(self, mins, maxs):
        self.extent = ConductorExtent(
            [self.xcent + mins[0], self.ycent + mins[1], self.zcent + mins[2]],
            [self.xcent + maxs[0], self.ycent + maxs[1], self.zcent + maxs[2]])

    def getextent(self):
        return self.extent

    def __add__(self, right):
        return AssemblyPlus(self, right)

    # Subclasses: number of intercepts per line and the native generator.
    def nicpt(self):
        raise NotImplementedError

    def fill_intercepts(self, icpt, fuzz):
        raise NotImplementedError


class AssemblyPlus(Assembly):
    """Union of two assemblies; installed part by part
    (generateconductors.py:2783-2786)."""

    def __init__(self, left, right):
        self.left = left
        self.right = right
        self.condid = left.condid
        self.voltage = left.voltage
        self.neumann = left.neumann
        el, er = left.getextent(), right.getextent()
        self.extent = ConductorExtent([min(a, b) for a, b in zip(el.mins, er.mins)],
                                      [max(a, b) for a, b in zip(el.maxs, er.maxs)])


class ZCylinder(Assembly):
    """Cylinder aligned with z (generateconductors.py:3305-)."""

    def __init__(self, radius, length=None, voltage=0., xcent=0., ycent=0., zcent=0.,
                 condid='next', zlower=None, zupper=None, neumann=0):
        if zlower is not None and zupper is not None:
            length = zupper - zlower
            zcent = 0.5 * (zlower + zupper)
        assert length is not None, "ZCylinder: either length or both zlower and zupper must be specified"
        Assembly.__init__(self, voltage, xcent, ycent, zcent, condid, neumann)
        self.radius = radius
        self.length = length
        self.createextent([-radius, -radius, -length / 2.], [+radius, +radius, +length / 2.])

    def nicpt(self):
        return 2, 2, 2

    def fill_intercepts(self, icpt, fuzz):
        N.call("wm_zcylinder_intercepts", self.radius, self.length, self.xcent, self.ycent,
               self.zcent, C.addressof(icpt), fuzz, 0)


class ZCylinderOut(Assembly):
    """Outside of a cylinder aligned with z (generateconductors.py:3494-3553)."""

    def __init__(self, radius, length=None, voltage=0., xcent=0., ycent=0., zcent=0.,
                 condid='next', zlower=None, zupper=None, neumann=0):
        if zlower is not None and zupper is not None:
            length = zupper - zlower
            zcent = 0.5 * (zlower + zupper)
        assert length is not None, "ZCylinderOut: either length or both zlower and zupper must be specified"
        Assembly.__init__(self, voltage, xcent, ycent, zcent, condid, neumann)
        self.radius = radius
        self.length = length
        self.createextent([-LARGEPOS, -LARGEPOS, -length / 2.],
                          [+LARGEPOS, +LARGEPOS, +length / 2.])

    def nicpt(self):
        return 4, 4, 2

    def fill_intercepts(self, icpt, fuzz):
        N.call("wm_zcylinder_intercepts", self.radius, self.length, self.xcent, self.ycent,
               self.zcent, C.addressof(icpt), fuzz, 1)


class ZSrfrvIn(Assembly):
    """Inside of a surface of revolution given by straight segments
    (generateconductors.py:4895-5018): the closed r-z outline
    rmin -> rofzdata -> rmin."""

    def __init__(self, zmin, zmax, rmin=0., rofzdata=None, zdata=None, voltage=0.,
                 xcent=0., ycent=0., zcent=0., condid='next', neumann=0):
        Assembly.__init__(self, voltage, xcent, ycent, zcent, condid, neumann)
        self.zmin, self.zmax, self.rmin = zmin, zmax, rmin
        self.rofzdata = list(rofzdata)
        self.zdata = list(zdata)
        rmax = max(self.rofzdata)
        self.createextent([-rmax, -rmax, zmin], [+rmax, +rmax, zmax])
        self.rsrf = [rmin] + self.rofzdata + [rmin, rmin]
        self.zsrf = [zmin] + self.zdata + [zmax, zmin]

    def nicpt(self):
        n = 2 * (len(self.rsrf) - 1)
        return n, n, n

    def fill_intercepts(self, icpt, fuzz):
        r = np.ascontiguousarray(self.rsrf, np.float64)
        z = np.ascontiguousarray(self.zsrf, np.float64)
        N.call("wm_zsrfrv_intercepts", len(r), r.ctypes.data, z.ctypes.data, self.xcent,
               self.ycent, self.zcent, C.addressof(icpt))


class ZAnnulus(ZSrfrvIn):
    """Annulus as a surface of revolution (generateconductors.py:5821-5855):
    z data in the lab frame with zcent = 0."""

    def __init__(self, rmin, rmax, length, voltage=0., xcent=0., ycent=0., zcent=0.,
                 condid='next', neumann=0):
        self.length = length
        self.rmax = rmax
        zmin = -length / 2. + zcent
        zmax = +length / 2. + zcent
        ZSrfrvIn.__init__(self, zmin, zmax, rmin=rmin, rofzdata=[rmax, rmax],
                          zdata=[zmin, zmax], voltage=voltage, xcent=xcent, ycent=ycent,
                          zcent=0., condid=condid, neumann=neumann)


def QuadRods(rodradius, rodcenter, rodlength, voltage, zcent=0., xcent=0., ycent=0.,
             condid='next'):
    """The four rods of an electrostatic quadrupole as ZCylinders: +voltage
    on the x rods and -voltage on the y rods (the rod part of Quadrupole,
    generateconductors.py:6021-)."""
    a = ZCylinder(rodradius, rodlength, voltage, xcent + rodcenter, ycent, zcent, condid)
    b = ZCylinder(rodradius, rodlength, voltage, xcent - rodcenter, ycent, zcent, condid)
    c = ZCylinder(rodradius, rodlength, -voltage, xcent, ycent + rodcenter, zcent, condid)
    d = ZCylinder(rodradius, rodlength, -voltage, xcent, ycent - rodcenter, zcent, condid)
    return a + b + c + d


def _nint(x):
    return int(np.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def getmglevels(nx, ny, nz, dx, dy, dz):
    """getmglevels / getnextcoarselevel3d (f3d_mgrid.F:3601-3845), serial:
    list of dicts nx, ny, nz, lx, ly, lz per level."""
    levels = []
    lx = ly = lz = 1.
    while True:
        levels.append(dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz))
        if nx < 4 or (ny < 4 and ny > 0) or nz < 4 or len(levels) == 101:
            break
        nxc, nyc, nzc = (nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2
        nxc += nxc % 2
        nyc += nyc % 2
        nzc += nzc % 2
        mn = min(min(dx, dy), dz)
        if dx > 4. / 3. * mn: nxc = nx
        if dy > 4. / 3. * mn: nyc = ny
        if dz > 4. / 3. * mn: nzc = nz
        dxc = dx * nx / nxc
        dyc = dxc if nyc == 0 else dy * ny / nyc
        dzc = dz * nz / nzc
        lx, ly, lz = lx * (dxc / dx), ly * (dyc / dy), lz * (dzc / dz)
        nx, ny, nz, dx, dy, dz = nxc, nyc, nzc, dxc, dyc, dzc
    return levels


class Grid:
    """Mesh description used to generate conductor data
    (generateconductors.py:2339-2660), serial, zbeam = 0."""

    def __init__(self, nx, ny, nz, xmmin, xmmax, ymmin, ymmax, zmmin, zmmax, zbeam=0.,
                 xmin=None, xmax=None, ymin=None, ymax=None, zmin=None, zmax=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.xmmin, self.xmmax = xmmin, xmmax
        self.ymmin, self.ymmax = ymmin, ymmax
        self.zmmin, self.zmmax = zmmin, zmmax
        self.zbeam = zbeam
        # the region conductor data is generated in: the mesh unless the
        # caller crops it (generateconductors.py:2421-2426)
        d = lambda v, default: default if v is None else v
        self.xmin, self.xmax = d(xmin, xmmin), d(xmax, xmmax)
        self.ymin, self.ymax = d(ymin, ymmin), d(ymax, ymmax)
        self.zmin, self.zmax = d(zmin, zmmin + zbeam), d(zmax, zmmax + zbeam)
        self.dx = (xmmax - xmmin) / nx
        self.dy = (ymmax - ymmin) / ny
        self.dz = (zmmax - zmmin) / nz
        self.levels = getmglevels(nx, ny, nz, self.dx, self.dy, self.dz)

    def getmeshsize(self, mglevel):
        L = self.levels[mglevel]
        return (self.dx * L["lx"], self.dy * L["ly"], self.dz * L["lz"],
                L["nx"], L["ny"], L["nz"])

    def getmeshnew(self, mglevel, extent):
        """Grid.getmeshnew (generateconductors.py:2592-2635)."""
        dx, dy, dz, nxl, nyl, nzl = self.getmeshsize(mglevel)
        zb = self.zbeam
        xmin = max(self.xmin, self.xmmin + 0 * dx)
        xmax = min(self.xmax, self.xmmin + (0 + nxl) * dx)
        ymin = max(self.ymin, self.ymmin + 0 * dy)
        ymax = min(self.ymax, self.ymmin + (0 + nyl) * dy)
        zmin = max(self.zmin, self.zmmin + 0 * dz + zb)
        zmax = min(self.zmax, self.zmmin + (0 + nzl) * dz + zb)
        lo, hi = extent.mins, extent.maxs
        if lo[0] - dx > xmin:
            xmin = int((lo[0] - dx - xmin) / dx) * dx + xmin
            if xmin >= xmax: xmin = xmax - 2 * dx
        if lo[1] - dy > ymin:
            ymin = int((lo[1] - dy - ymin) / dy) * dy + ymin
            if ymin >= ymax: ymin = ymax - 2 * dy
        if lo[2] - dz > zmin:
            zmin = int((lo[2] - dz - zmin) / dz) * dz + zmin
            if zmin >= zmax: zmin = zmax - 2 * dz
        if hi[0] + dx < xmax:
            xmax = int((hi[0] + dx - xmin) / dx) * dx + xmin
            if xmax < hi[0] + dx: xmax += dx
            if xmax <= xmin: xmax = xmin + 2 * dx
        if hi[1] + dy < ymax:
            ymax = int((hi[1] + dy - ymin) / dy) * dy + ymin
            if ymax < hi[1] + dy: ymax += dy
            if ymax <= ymin: ymax = ymin + 2 * dy
        if hi[2] + dz < zmax:
            zmax = int((hi[2] + dz - zmin) / dz) * dz + zmin
            if zmax < hi[2] + dz: zmax += dz
            if zmax <= zmin: zmax = zmin + 2 * dz
        return (xmin, ymin, zmin, dx, dy, dz,
                _nint((xmax - xmin) / dx), _nint((ymax - ymin) / dy),
                _nint((zmax - zmin) / dz),
                _nint((xmin - (self.xmmin + 0 * dx)) / dx),
                _nint((ymin - (self.ymmin + 0 * dy)) / dy),
                _nint((zmin - (self.zmmin + 0 * dz + zb)) / dz))

    def checkoverlap(self, mglevel, extent):
        """Grid.checkoverlap (generateconductors.py:2637-2660), serial."""
        dx, dy, dz = self.getmeshsize(mglevel)[:3]
        xmin = max(extent.mins[0], self.xmin)
        ymin = max(extent.mins[1], self.ymin)
        zmin = max(extent.mins[2], self.zmin)
        xmax = min(extent.maxs[0], self.xmax)
        ymax = min(extent.maxs[1], self.ymax)
        zmax = min(extent.maxs[2], self.zmax)
        return not (zmin - dz > zmax or xmin - dx > xmax or ymin - dy > ymax)

    def intercepts(self, a, mglevel):
        """One object's intercepts on one level (Assembly.gridintercepts +
        GridIntercepts.__init__, generateconductors.py:333-340, 1815-1849).
        Returns the WmIntercepts struct and the numpy arrays it points to."""
        xmin, ymin, zmin, dx, dy, dz, nx, ny, nz, ix, iy, iz = self.getmeshnew(mglevel, a.extent)
        nxi, nyi, nzi = a.nicpt()
        s = N.WmIntercepts()
        s.xmmin, s.ymmin, s.zmmin, s.dx, s.dy, s.dz = xmin, ymin, zmin, dx, dy, dz
        s.nx, s.ny, s.nz, s.ix, s.iy, s.iz = nx, ny, nz, ix, iy, iz
        s.mglevel = mglevel
        s.nxicpt, s.nyicpt, s.nzicpt = nxi, nyi, nzi
        keep = {}
        shapes = dict(x=(nz + 1, ny + 1, nxi), y=(nz + 1, nx + 1, nyi), z=(ny + 1, nx + 1, nzi))
        for k, shp in shapes.items():
            shp = tuple(max(v, 0) for v in shp)
            keep[k] = np.zeros(shp)
            keep[k + "v"] = np.full(shp, float(a.voltage))
            keep[k + "c"] = np.full(shp, int(a.condid), np.int64)
            setattr(s, k + "intercepts", keep[k].ctypes.data_as(N._pd))
            setattr(s, k + "voltages", keep[k + "v"].ctypes.data_as(N._pd))
            setattr(s, k + "condids", keep[k + "c"].ctypes.data_as(N._pi))
        fuzz = 1.e-13
        if nx >= 0 and ny >= 0 and nz >= 0:
            a.fill_intercepts(s, fuzz)
        return s, keep


class ConductorData:
    """Installed conductors of one solver: the counterpart of ConductorType
    (pywarp90/f3d.v:105-164), held by libwarpmi on the host."""

    def __init__(self, nx, ny, nz, dx, dy, dz):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = dx, dy, dz
        h = C.c_void_p()
        N.call("wm_condlist_create", C.byref(h))
        self.h = h
        self.checked = False

    def __del__(self):
        try:
            if self.h:
                N.lib().wm_condlist_destroy(self.h)
        except Exception:
            pass

    def add_intercepts(self, icpt, dfill, fuzz=1.e-13):
        N.call("wm_conductordelfromintercepts", C.addressof(icpt), self.h, dfill, fuzz)
        self.checked = False

    def check(self):
        """checkconductors: levels, clean-up and sort by level."""
        N.call("wm_checkconductors", self.h, self.nx, self.ny, self.nz, self.dx, self.dy,
               self.dz)
        self.checked = True

    def view(self):
        if not self.checked:
            self.check()
        v = N.WmConductorLists()
        N.call("wm_condlist_view", self.h, C.addressof(v))
        return v

    def lists(self):
        """Copies of the sorted lists as numpy arrays."""
        v = self.view()
        L = v.levels

        def arr(p, n, w, dt):
            if n == 0:
                return np.zeros((0, w) if w > 1 else (0,), dt)
            a = np.ctypeslib.as_array(p, shape=(n * w,)).copy()
            return a.reshape(n, w) if w > 1 else a
        out = dict(levels=L,
                   int_indx=arr(v.int_indx, v.int_n, 3, np.int64),
                   int_volt=arr(v.int_volt, v.int_n, 1, np.float64),
                   int_istart=arr(v.int_istart, L + 1, 1, np.int64) if v.int_n else np.zeros(L + 1, np.int64))
        for p in range(2):
            n = v.sub_n[p]
            out["sub%d_indx" % p] = arr(v.sub_indx[p], n, 3, np.int64)
            out["sub%d_dels" % p] = arr(v.sub_dels[p], n, 6, np.float64)
            out["sub%d_volt" % p] = arr(v.sub_volt[p], n, 6, np.float64)
            out["sub%d_istart" % p] = (arr(v.sub_istart[p], L + 1, 1, np.int64) if n
                                       else np.zeros(L + 1, np.int64))
        return out


def installconductor(a, grid, conductors, dfill=2.0):
    """installconductor -> Grid.getdatanew -> GridIntercepts.installlist
    (generateconductors.py:135-215, 2750-2808, 1851-1948), serial."""
    if isinstance(a, (list, tuple)):
        for c in a:
            installconductor(c, grid, conductors, dfill)
        return
    if dfill is None:
        dfill = 2.0
    if not grid.checkoverlap(len(grid.levels) - 1, a.getextent()):
        return
    if isinstance(a, AssemblyPlus):
        installconductor(a.left, grid, conductors, dfill)
        installconductor(a.right, grid, conductors, dfill)
        return
    for mglevel in range(len(grid.levels)):
        icpt, keep = grid.intercepts(a, mglevel)
        conductors.add_intercepts(icpt, dfill)
        del keep
