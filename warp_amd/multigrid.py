"""MultiGrid3D, the drop-in for scripts/multigrid.py's registered 3-D
multigrid field solver, running on an MI355X through libwarpmi.

Interface as in the reference (multigrid.py:19-953): construct with the
mesh and f3d multigrid inputs, ``registersolver(solver)``, install
conductors with ``installconductor``, then per step ``loadrho()``,
``solve()``, ``fetche()``.  ``phi``, ``rho`` and ``selfe`` are FP64 torch
tensors on the GPU in the reference's Fortran layout: phi is
(-1:nx+1,-1:ny+1,-1:nz+1), i.e. a C-contiguous tensor of shape
(nz+3, ny+3, nx+3); rho is (nz+1, ny+1, nx+1).  After ``solve`` the solver
holds ``mgiters`` and ``mgerror``; non-convergence is reported, not raised
(f3d_mgrid.F:460-467).

The hot loop never falls back to the CPU: if libwarpmi or a HIP device is
missing, construction fails.
"""
import ctypes as C
import math
import sys
import time

import torch

from . import _native as N
from .fieldsolver import FieldSolver
from .generateconductors import ConductorData, Grid, installconductor as _install
from .globals import eps0, f3d, top, w3d

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}


class MultiGrid3D(FieldSolver):
    __f3dinputs__ = ['gridmode', 'mgparam', 'downpasses', 'uppasses', 'mgmaxiters', 'mgtol',
                     'mgmaxlevels', 'mgform', 'mgverbose', 'mgntverbose', 'lcndbndy',
                     'icndbndy', 'laddconductor', 'lprecalccoeffs']

    def __init__(self, precision="fp64", device=None, sorted_deposit=False, guess="previous",
                 learned_weights=None, sorform=0, **kw):
        FieldSolver.__init__(self, kwdict=kw)
        self.processdefaultsfrompackage(MultiGrid3D.__f3dinputs__, f3d, kw)
        assert len(kw.keys()) == 0, "Bad keyword arguemnts %s" % list(kw.keys())
        if self.lprecalccoeffs:
            raise NotImplementedError("lprecalccoeffs: the coefficients are always derived on the fly")
        if precision not in PRECISIONS:
            raise ValueError("precision must be one of %s" % list(PRECISIONS))
        self.precision = precision
        self.sorted_deposit = sorted_deposit
        if guess not in ("previous", "linear", "learned"):
            raise ValueError("guess must be 'previous' (the reference's), 'linear' or 'learned'")
        # initial guess of each solve: the previous phi (the reference), its
        # linear extrapolation from the two previous solves, or that plus the
        # learned correction (warp_amd.learned; falls back to the
        # extrapolation when the network's residual is not smaller)
        self.guess = guess
        self._phiprev = None
        self._learned = None
        self._learned_weights = learned_weights
        if not torch.cuda.is_available():
            raise RuntimeError("MultiGrid3D needs a HIP device (torch.cuda is not available)")
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("MultiGrid3D needs a HIP device, got %s" % self.device)
        N.lib()
        self.solvergeom = "XYZgeom"
        self.mgiters = 0
        self.mgiters_history = []  # mgiters of every solve, in order
        self.mgerror = 0.
        self.lwithselfe = 0
        self._phi = torch.zeros((self.nz + 3, self.ny + 3, self.nx + 3), dtype=torch.float64,
                                device=self.device)
        self._rho = torch.zeros((self.nz + 1, self.ny + 1, self.nx + 1), dtype=torch.float64,
                                device=self.device)
        self._selfe = None
        self._work = None
        self._nlost = None
        self._cf = None            # conductor field data for f3d.lcorrectede
        self._cf_key = None
        self._cond_gen = 0         # counts uploads of conductor lists
        self._phi_token = None     # (version, address) of phi after the last solve
        self._planes = None        # plane charges of gtlchg
        self._esework = None       # column sums of getese, then ese itself
        b = (C.c_int64 * 6)(*self.bounds)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            N.call("wm_mg3d_create", C.byref(h), self.nx, self.ny, self.nz, self.dx, self.dy,
                   self.dz, b, self.mgmaxlevels)
        self._h = h
        # SOR pass form: 0 two half sweeps (one launch per colour), 1 both
        # colours in one launch where the level allows it, 2 that on the
        # finest level only; the same bits
        if sorform not in (0, 1, 2):
            raise ValueError("sorform must be 0 (two half sweeps), 1 (fused pass) or 2 (fused "
                             "pass on the finest level)")
        self.sorform = sorform
        N.call("wm_mg3d_set_sor_form", h, sorform)
        self.initializeconductors()
        self.timers = dict(deposit=0., solve=0., fetche=0.)

    def __del__(self):
        try:
            self._drop_condfield()
            if getattr(self, "_h", None):
                N.lib().wm_mg3d_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # ------------------------------------------------------------------
    # arrays (multigrid.py:317-421)
    @property
    def phi(self):
        "Electrostatic potential array, including guard cells"
        return self._phi

    @phi.setter
    def phi(self, value):
        self._phi[...] = value

    @property
    def rho(self):
        "Charge density array"
        return self._rho

    @rho.setter
    def rho(self, value):
        self._rho[...] = value

    @property
    def selfe(self):
        "Electric field array (3, nx+1, ny+1, nz+1) in Fortran order"
        self.calcselfe()
        return self._selfe

    def getphi(self):
        'Returns the phi array without the guard cells'
        return self._phi[1:-1, 1:-1, 1:-1]

    def getrho(self):
        'Returns the rho array without the guard cells'
        return self._rho

    def getselfe(self, recalculate=None, lzero=True):
        'Returns the E field array without the guard cells'
        self.calcselfe(lzero=lzero)
        return self._selfe

    def _gridfield(self):
        """selfe of the current phi (recomputed after each solve) for the
        efetch=3 gathers."""
        key = (len(self.mgiters_history), bool(f3d.lcorrectede))
        if self._selfe is None or getattr(self, "_selfe_key", None) != key:
            self.calcselfe(lzero=True)
            if f3d.lcorrectede:
                # E at the conductor nodes (multigrid.py:638-646)
                N.call("wm_condfield_fixefield", self._condfield(), self._selfe.data_ptr(),
                       N.stream_of(self.device))
            self._selfe_key = key
        return self._selfe

    # ------------------------------------------------------------------
    # conductor-corrected E (f3d.lcorrectede; multigrid.py:528-546, 602-646)
    def _drop_condfield(self):
        cf = getattr(self, "_cf", None)
        if cf:
            N.lib().wm_condfield_destroy(cf)
        self._cf = None
        self._cf_key = None

    def _condfield(self):
        """setupconductorfielddata (f3d_mgrid.F:4581-4633): the node
        classification of the installed conductors on this grid, kept
        until the conductors change."""
        if self.conductordatalist:
            obj = self.getconductorobject()
        else:
            obj = None
        key = self._cond_gen
        if self._cf is None or self._cf_key != key:
            self._drop_condfield()
            view = obj.view() if obj is not None else None
            h = C.c_void_p()
            with torch.cuda.device(self.device):
                N.call("wm_condfield_create", C.byref(h), self.nx, self.ny, self.nz,
                       C.addressof(view) if view is not None else None)
            self._cf = h
            self._cf_key = key
            self._cf_phi_key = None
        return self._cf

    def _condfield_efield(self):
        """getefieldatconductorsubgrid (f3d_mgrid.F:4635-4735) of the
        current phi, once per solve."""
        cf = self._condfield()
        key = len(self.mgiters_history)
        if getattr(self, "_cf_phi_key", None) != key:
            N.call("wm_condfield_efield", cf, self._phi.data_ptr(), self.dx, self.dy, self.dz,
                   N.stream_of(self.device))
            self._cf_phi_key = key
        return cf

    def calcselfe(self, recalculate=None, lzero=True):
        """getselfe3d (w3d.F:6023-6109): E = -grad phi by centred
        differences, stored (nz+1, ny+1, nx+1, 3)."""
        if self._selfe is None:
            self._selfe = torch.zeros((self.nz + 1, self.ny + 1, self.nx + 1, 3),
                                      dtype=torch.float64, device=self.device)
        N.call("wm_getselfe3d", self._phi.data_ptr(), self.nx, self.ny, self.nz,
               self._selfe.data_ptr(), self.dx, self.dy, self.dz, int(lzero),
               N.stream_of(self.device))

    # ------------------------------------------------------------------
    # conductors (multigrid.py:121-250, 743-825)
    def initializeconductors(self):
        self.conductordatalist = []
        self._conductors = None
        self._installedconductors = []
        self._cond_gen = getattr(self, "_cond_gen", 0) + 1

    def installconductor(self, conductor, xmin=None, xmax=None, ymin=None, ymax=None,
                         zmin=None, zmax=None, dfill=None):
        'Install the conductor into the field solver'
        # as the reference, only recorded here; the data is generated when
        # it is needed (getconductorobject), within xmin..zmax when given
        self.conductordatalist.append((conductor, xmin, xmax, ymin, ymax, zmin, zmax, dfill))

    def uninstallconductor(self, conductor):
        n = len(self.conductordatalist)
        self.conductordatalist = [c for c in self.conductordatalist if c[0] is not conductor]
        if len(self.conductordatalist) == n:
            raise Exception('conductor was not found')
        self.clearconductors()

    def clearconductors(self):
        "Clear out the conductor data"
        self._conductors = None
        self._installedconductors = []
        self._drop_condfield()
        self._cond_gen += 1
        N.call("wm_mg3d_set_conductors", self._h, None)

    def hasconductors(self):
        return len(self.conductordatalist) > 0

    def getconductorobject(self):
        """Installs conductors not yet installed, uploads the lists to the
        GPU solver and returns the ConductorData."""
        pending = [c for c in self.conductordatalist
                   if not any(c[0] is d for d in self._installedconductors)]
        if self._conductors is None or pending:
            if self._conductors is None:
                self._conductors = ConductorData(self.nx, self.ny, self.nz, self.dx, self.dy,
                                                 self.dz)
                pending = list(self.conductordatalist)
                self._installedconductors = []
            for conductor, xmin, xmax, ymin, ymax, zmin, zmax, dfill in pending:
                grid = Grid(self.nx, self.ny, self.nz, self.xmmin, self.xmmax, self.ymmin,
                            self.ymmax, self.zmmin, self.zmmax, xmin=xmin, xmax=xmax,
                            ymin=ymin, ymax=ymax, zmin=zmin, zmax=zmax)
                _install(conductor, grid, self._conductors, dfill if dfill is not None else 2.0)
                self._installedconductors.append(conductor)
            view = self._conductors.view()
            N.call("wm_mg3d_set_conductors", self._h, C.addressof(view))
            self._cond_gen += 1
        return self._conductors

    # ------------------------------------------------------------------
    # particles -> rho (fieldsolver.py:1333-1422; multigrid.py:472-510)
    def zerosourcep(self):
        self._rho.zero_()

    def loadrho(self, lzero=None, lfinalize_rho=None, pgroups=None, **kw):
        if not self.ldosolve and lzero:
            return
        if lzero is None:
            lzero = w3d.lzerorhofsapi
        if lfinalize_rho is None:
            lfinalize_rho = w3d.lfinalizerhofsapi
        t0 = time.perf_counter()
        if lzero:
            self.zerosourcep()
        if pgroups is None:
            pgroups = [top.pgroup] if top.pgroup is not None else []
        jslist = kw.get('jslist', None)
        for pg in pgroups:
            for js in (jslist if jslist is not None else range(pg.ns)):
                if pg.nps[js] == 0:
                    continue
                self.setsourcep(js, pg, top.zgrid)
        self.sourcepfinalized = False
        if lzero and lfinalize_rho:
            self.finalizesourcep()
        self.timers["deposit"] += time.perf_counter() - t0

    def setsourcep(self, js, pgroup, zgrid):
        sl = pgroup.species_slice(js)
        x, y, z = pgroup.xp[sl], pgroup.yp[sl], pgroup.zp[sl]
        q = pgroup.sq[js]
        w = pgroup.sw[js] * pgroup.dtscale[js]
        self.setsourcepatposition(x, y, z, zgrid, q, w)

    def setsourcepatposition(self, x, y, z, zgrid, q, w):
        """setrho3d (w3d.F:7199-7301) on the GPU."""
        n = int(x.numel())
        if n == 0:
            return
        s = N.stream_of(self.device)
        if self.sorted_deposit:
            need = N.lib().wm_setrho3d_sorted_worksize(n, self.nx, self.ny, self.nz)
            if self._work is None or self._work.numel() < need:
                self._work = torch.empty(int(need), dtype=torch.uint8, device=self.device)
            N.call("wm_setrho3d_sorted", self._rho.data_ptr(), n, x.data_ptr(), y.data_ptr(),
                   z.data_ptr(), zgrid, q, w, self.nx, self.ny, self.nz, self.dx, self.dy,
                   self.dz, self.xmmin, self.ymmin, self.zmmin, self._work.data_ptr(), s)
        else:
            N.call("wm_setrho3d", self._rho.data_ptr(), n, x.data_ptr(), y.data_ptr(),
                   z.data_ptr(), zgrid, q, w, self.nx, self.ny, self.nz, self.dx, self.dy,
                   self.dz, self.xmmin, self.ymmin, self.zmmin, None, s)

    def finalizesourcep(self):
        if self.sourcepfinalized:
            return
        self.sourcepfinalized = True
        self.applysourceboundaryconditions()

    def applysourceboundaryconditions(self):
        """applyrhoboundaryconditions3d (w3d.F:2660-2750)."""
        b = (C.c_int64 * 6)(*self.bounds)
        N.call("wm_applyrhoboundaryconditions3d", self.nx, self.ny, self.nz,
               self._rho.data_ptr(), b, N.stream_of(self.device))

    # ------------------------------------------------------------------
    # the solve (fieldsolver.py:1575-1594; multigrid.py:846-914)
    def solve(self, iwhich=0):
        if not self.ldosolve:
            return
        self.finalizesourcep()
        self.dosolve(iwhich)

    def find_mgparam(self, lsavephi=False, resetpasses=1):
        """Optimise mgparam and the pass counts (multigrid.py:827-831,
        scripts/find_mgparam.py)."""
        from .find_mgparam import find_mgparam
        find_mgparam(self, lsavephi=lsavephi, resetpasses=resetpasses)

    def getmgverbose(self):
        if self.mgntverbose < 0:
            return 0
        if self.mgntverbose > 1:
            return 1 if (top.it % self.mgntverbose) == 0 else 0
        return self.mgverbose

    def dosolve(self, iwhich=0, *args):
        if not self.l_internal_dosolve:
            return
        if iwhich == 1:
            return
        t0 = time.perf_counter()
        if self.conductordatalist:
            self.getconductorobject()
        if self.guess == "linear":
            self._extrapolate_guess()
        elif self.guess == "learned":
            self._learned_guess()
        it = C.c_int64(0)
        err = C.c_double(0.)
        # phi untouched since the last solve returned it (torch's version
        # counter counts every in-place change made through the tensor or
        # its views): the solve may start from its own copy
        tok = self._phi_version()
        if tok is not None and self._phi_token == tok:
            N.call("wm_mg3d_reuse_phi", self._h)
        N.call("wm_mg3d_solve", self._h, self._phi.data_ptr(), self._rho.data_ptr(),
               self.mgparam, self.mgform, self.mgmaxiters, self.mgtol, self.downpasses,
               self.uppasses, int(self.lcndbndy), self.icndbndy,
               PRECISIONS[self.precision], C.byref(it), C.byref(err),
               N.stream_of(self.device))
        self._phi_token = self._phi_version()
        self.mgiters = it.value
        self.mgiters_history.append(self.mgiters)
        self.mgerror = err.value
        f3d.mgiters, f3d.mgerror = self.mgiters, self.mgerror
        if f3d.lcorrectede:
            self._condfield_efield()
        if top.efetch == 3:
            # the grid field the efetch=3 gather interpolates, computed once
            # per solve (setfieldpforparticles -> calcselfep, multigrid.py:617-629)
            self._gridfield()
        mgverbose = self.getmgverbose()
        if mgverbose >= 1 or (mgverbose >= 0 and self.mgerror > self.mgtol):
            if self.mgerror > self.mgtol:
                print("Multigrid: Maximum number of iterations reached", file=sys.stderr)
            print("Multigrid: Error converged to %11.3e in %5d v-cycles"
                  % (self.mgerror, self.mgiters), file=sys.stderr)
        self.timers["solve"] += time.perf_counter() - t0

    def last_solve_reused_phi(self):
        """True when the last solve started from the solver's own copy of
        the phi the solve before it returned (phi untouched in between)."""
        r = C.c_int64(0)
        N.call("wm_mg3d_last_reused", self._h, C.byref(r))
        return bool(r.value)

    def tail_start_level(self):
        """The first level of the V-cycle's coarse end that the last solve
        ran as one launch (the levels' arrays in LDS), or -1."""
        r = C.c_int64(0)
        N.call("wm_mg3d_tail_start", self._h, C.byref(r))
        return int(r.value)

    def _phi_version(self):
        """(version, address) of phi, or None where torch keeps no version
        counter (tensors made under torch.inference_mode)."""
        try:
            return (self._phi._version, self._phi.data_ptr())
        except RuntimeError:
            return None

    def learned_guess_state(self):
        """The LearnedGuess of guess="learned" (loaded on first use): its
        network, and the counts of solves that used the network's guess
        (applied) and that fell back to the extrapolation (fallbacks)."""
        if self._learned is None:
            from . import learned as LG
            path = self._learned_weights or LG.default_weights()
            net, cphys, meta = LG.load_weights(path, self.device)
            self._learned = LG.LearnedGuess(self, net, cphys)
            self._learned.meta = meta
        return self._learned

    def _learned_guess(self):
        """Linear extrapolation plus the network's correction of it, kept
        only when it lowers the largest residual (else the extrapolation
        alone: the plain method is the fallback)."""
        from .learned import residual_max
        lg = self.learned_guess_state()
        lg.observe(self._rho)
        extrapolated = self._extrapolate_guess()
        if not (extrapolated and lg.ready()):
            return
        base = self._phi.clone()
        r_lin = residual_max(self, base)
        lg.correct(self._phi)
        r_net = residual_max(self, self._phi)
        if r_net < r_lin:
            lg.applied += 1
        else:
            self._phi.copy_(base)
            lg.fallbacks += 1

    def _extrapolate_guess(self):
        """phi <- 2 phi_(n-1) - phi_(n-2) before a solve (the previous phi
        alone while fewer than two solves have run); keeps phi_(n-1) for the
        next one."""
        if not self.mgiters_history:
            return False  # phi is not the result of a solve yet
        if self._phiprev is None:
            self._phiprev = self._phi.clone()
            return False
        N.call("wm_phi_extrapolate", self._phi.data_ptr(), self._phiprev.data_ptr(),
               self._phi.numel(), N.stream_of(self.device))
        self._phi_token = None  # phi changed outside torch's view
        return True

    # ------------------------------------------------------------------
    # field diagnostics of special steps (multigrid.py:433-467)
    def gtlchg(self):
        """Line charge of rho into top.linechg (gtlchg3dfromrho,
        dw3d.F:229-320): a float64 tensor on the device with the
        top.nzzarr+1 points of the Z_arrays mesh.  The Z_arrays mesh
        defaults to the field mesh (w3d.F:337-341)."""
        if not w3d.lgtlchg3d:
            return
        if top.nzzarr == 0:
            top.nzzarr = self.nz
        if top.zzmin == 0.:
            top.zzmin = self.zmmin
        if top.zzmax == 0.:
            top.zzmax = self.zmmax
        top.dzz = (top.zzmax - top.zzmin) / top.nzzarr
        lc = top.linechg
        if not (isinstance(lc, torch.Tensor) and lc.device == self.device
                and lc.numel() == top.nzzarr + 1):
            lc = top.linechg = torch.zeros(top.nzzarr + 1, dtype=torch.float64,
                                           device=self.device)
        if self._planes is None or self._planes.numel() != self.nz + 1:
            self._planes = torch.empty(self.nz + 1, dtype=torch.float64, device=self.device)
        N.call("wm_gtlchg3d", self._rho.data_ptr(), self.nx, self.ny, self.nz, self.dx, self.dy,
               self.dz, top.zgrid, self.zmminlocal, top.zzmin, top.zbeam, top.dzz, top.nzzarr,
               int(self.l2symtry), int(self.l4symtry), self._planes.data_ptr(), lc.data_ptr(),
               N.stream_of(self.device))

    def getese(self):
        """Electrostatic energy sum(rho*phi)*dx*dy*dz/2 into top.ese
        (getese3dfromrhophi, dw3d.F:708-768)."""
        need = int(N.lib().wm_getese3d_worksize(self.nx, self.ny))
        ncol = need // 8
        if self._esework is None or self._esework.numel() < ncol + 1:
            self._esework = torch.empty(ncol + 1, dtype=torch.float64, device=self.device)
        out = self._esework[ncol:ncol + 1]
        N.call("wm_getese3d", self._rho.data_ptr(), self._phi.data_ptr(), self.nx, self.ny,
               self.nz, self.dx, self.dy, self.dz, int(self.l4symtry), int(self.l2symtry),
               self._esework.data_ptr(), out.data_ptr(), N.stream_of(self.device))
        top.ese = float(out.item())

    def getresidual(self):
        """residual3d (f3d_mgrid.F:3367-3598) of the current phi and rho,
        scaled by mgparam/(eps0*2*(dxsqi+dysqi+dzsqi)) as the reference's
        multigrid.py:940-951."""
        dxsqi, dysqi, dzsqi = 1. / self.dx ** 2, 1. / self.dy ** 2, 1. / self.dz ** 2
        reps0c = self.mgparam / (eps0 * 2. * (dxsqi + dysqi + dzsqi))
        rho = self._rho * reps0c
        res = torch.zeros_like(self._phi)
        if self.conductordatalist:
            self.getconductorobject()
        N.call("wm_mg3d_residual", self._h, 0, self._phi.data_ptr(), rho.data_ptr(),
               res.data_ptr(), self.mgparam, self.mgform, 0, int(self.lcndbndy),
               self.icndbndy, N.stream_of(self.device))
        return res

    # ------------------------------------------------------------------
    # fields -> particles (fieldsolver.py:1434-1567; multigrid.py:521-575)
    def fetche(self, *args, **kw):
        'Fetches the field, uses arrays from the w3d FieldSolveAPI globals'
        if w3d.npfsapi == 0:
            return
        t0 = time.perf_counter()
        if w3d.xfsapi is not None:
            x, y, z = w3d.xfsapi, w3d.yfsapi, w3d.zfsapi
            ex, ey, ez = w3d.exfsapi, w3d.eyfsapi, w3d.ezfsapi
        else:
            pg = w3d.pgroupfsapi
            i = w3d.ipminfsapi - 1
            sl = slice(i, i + w3d.npfsapi)
            x, y, z = pg.xp[sl], pg.yp[sl], pg.zp[sl]
            ex, ey, ez = pg.ex[sl], pg.ey[sl], pg.ez[sl]
        self.fetchfieldfrompositions(x, y, z, ex, ey, ez)
        self.timers["fetche"] += time.perf_counter() - t0

    def fetchfieldfrompositions(self, x, y, z, ex, ey, ez, *args, **kw):
        """sete3d (w3d.F:3656-3777): adds E at the positions to ex, ey, ez,
        from the centred differences of phi (sete_from_phi_linear3d,
        3963-4107) or, with top.efetch == 3, by interpolation of the grid
        field selfe (sete_from_e_linear3d, 4362-4460); with f3d.lcorrectede
        by sete3dwithconductor (f3d_mgrid.F:4774-5010), as multigrid.py:528-546."""
        n = int(x.numel())
        if n == 0:
            return
        for t in (x, y, z, ex, ey, ez):
            if not t.is_contiguous():
                raise ValueError("fetche arrays must be contiguous")
        if top.efetch not in (1, 3):
            raise NotImplementedError("efetch=%s: only 1 (from phi) and 3 (from the grid E)"
                                      % top.efetch)
        if f3d.lcorrectede:
            cf = self._condfield_efield()
            selfe = self._gridfield().data_ptr() if top.efetch == 3 else None
            N.call("wm_sete3d_conductor", cf, self._phi.data_ptr(), selfe, top.efetch, n,
                   x.data_ptr(), y.data_ptr(), z.data_ptr(), top.zgrid, self.xmmin, self.ymmin,
                   self.zmmin, self.dx, self.dy, self.dz, ex.data_ptr(), ey.data_ptr(),
                   ez.data_ptr(), N.stream_of(self.device))
            return
        if top.efetch == 3:
            N.call("wm_sete3d_from_e", self._gridfield().data_ptr(), n, x.data_ptr(), y.data_ptr(),
                   z.data_ptr(), top.zgrid, self.xmmin, self.ymmin, self.zmmin, self.dx,
                   self.dy, self.dz, self.nx, self.ny, self.nz, ex.data_ptr(), ey.data_ptr(),
                   ez.data_ptr(), N.stream_of(self.device))
            return
        N.call("wm_sete3d", self._phi.data_ptr(), n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
               top.zgrid, self.xmmin, self.ymmin, self.zmmin, self.dx, self.dy, self.dz,
               self.nx, self.ny, self.nz, ex.data_ptr(), ey.data_ptr(), ez.data_ptr(),
               N.stream_of(self.device))

    def fetchphi(self, *args, **kw):
        'Fetches the potential at w3d.xfsapi.. into w3d.phifsapi'
        if w3d.npfsapi == 0:
            return
        self.fetchpotentialfrompositions(w3d.xfsapi, w3d.yfsapi, w3d.zfsapi, w3d.phifsapi)

    def fetchpotentialfrompositions(self, x, y, z, phi):
        """getgrid3d (pywarp90/dtop.F:2818-2920) over the potential
        including its guard cells (multigrid.py:560-575): trilinear
        interpolation into phi at the positions inside the guarded mesh,
        phi unchanged elsewhere."""
        n = int(x.numel())
        if n == 0:
            return
        for t in (x, y, z, phi):
            if not (isinstance(t, torch.Tensor) and t.is_contiguous() and t.device == self.device
                    and t.dtype == torch.float64):
                raise ValueError("fetchphi arrays must be contiguous float64 tensors on %s"
                                 % self.device)
        dx, dy, dz = self.dx, self.dy, self.dz
        nxp, nyp, nzp = self.nx + 2, self.ny + 2, self.nz + 2
        xmin, xmax = self.xmmin - dx, self.xmmax + dx
        ymin, ymax = self.ymmin - dy, self.ymmax + dy
        zmin, zmax = self.zmmin - dz + top.zgrid, self.zmmax + dz + top.zgrid
        N.call("wm_getgrid3d", n, x.data_ptr(), y.data_ptr(), z.data_ptr(), phi.data_ptr(),
               nxp, nyp, nzp, self._phi.data_ptr(), xmin, xmax, ymin, ymax, zmin, zmax,
               int(self.l2symtry), int(self.l4symtry), N.stream_of(self.device))

    # fused fullv advance used by the stepper
    _CENTERS = {"fullv": 0, "halfv": 1, "synchv": 2}

    def gather_push(self, pg, js, dt, center="fullv"):
        """fetche and padvnc3d's kicks and drift for species js in one pass
        (B = 0, w3d.F:1575-2330): "fullv" epush(dt/2) twice + xpush(dt),
        "halfv" epush(dt/2) + xpush(dt), "synchv" epush(dt/2).  E as
        fetchfieldfrompositions takes it (top.efetch 1 or 3, f3d.lcorrectede)."""
        sl = pg.species_slice(js)
        n = sl.stop - sl.start
        if n == 0:
            return
        col = lambda c: getattr(pg, c)[sl]
        t0 = time.perf_counter()
        if self._nlost is None:
            self._nlost = torch.zeros(1, dtype=torch.int64, device=self.device)
        if top.efetch not in (1, 3):
            raise NotImplementedError("efetch=%s: only 1 (from phi) and 3 (from the grid E)"
                                      % top.efetch)
        if f3d.lcorrectede:
            cf = self._condfield_efield()
            selfe = self._gridfield().data_ptr() if top.efetch == 3 else None
            N.call("wm_padvnc3d_conductor", cf, self._phi.data_ptr(), selfe, top.efetch, n,
                   col("xp").data_ptr(), col("yp").data_ptr(), col("zp").data_ptr(),
                   col("uxp").data_ptr(), col("uyp").data_ptr(), col("uzp").data_ptr(),
                   col("gaminv").data_ptr(), col("ex").data_ptr(), col("ey").data_ptr(),
                   col("ez").data_ptr(), top.zgrid, self.xmmin, self.ymmin, self.zmmin, self.dx,
                   self.dy, self.dz, pg.sq[js], pg.sm[js], dt, *self._particle_extent(),
                   self._nlost.data_ptr(), self._CENTERS[center],
                   N.stream_of(self.device))
            self.timers["fetche"] += time.perf_counter() - t0
            return
        if top.efetch == 3:
            fn, grid = "wm_padvnc3d_selfe", self._gridfield()
        else:
            fn, grid = "wm_padvnc3d", self._phi
        N.call(fn, grid.data_ptr(), n, col("xp").data_ptr(),
               col("yp").data_ptr(), col("zp").data_ptr(), col("uxp").data_ptr(),
               col("uyp").data_ptr(), col("uzp").data_ptr(), col("gaminv").data_ptr(),
               col("ex").data_ptr(), col("ey").data_ptr(), col("ez").data_ptr(), top.zgrid,
               self.xmmin, self.ymmin, self.zmmin, self.dx, self.dy, self.dz, self.nx,
               self.ny, self.nz, pg.sq[js], pg.sm[js], dt, *self._particle_extent(),
               self._nlost.data_ptr(), self._CENTERS[center], N.stream_of(self.device))
        self.timers["fetche"] += time.perf_counter() - t0

    def _particle_extent(self):
        """Upper extents of the absorbing particle boundaries:
        getparticleextent's xmmin+nx*dx, ... (w3d.F:11344-11356)."""
        return (self.xmmin + self.nx * self.dx, self.ymmin + self.ny * self.dy,
                self.zmmin + self.nz * self.dz)

    def take_lost_count(self):
        """Particles counted outside the mesh by gather_push since the last
        call (reads one integer from the device, then resets it)."""
        if self._nlost is None:
            return None
        n = int(self._nlost.item())
        if n:
            self._nlost.zero_()
        return n

    def sortparticles(self, pgroup):
        """particlesortxyzwithcopy on this solver's mesh (top.F:5095)."""
        from .particles import particlesortxyzwithcopy
        particlesortxyzwithcopy(pgroup, self.dx, self.dy, self.dz, self.xmmin, self.ymmin,
                                self.zmmin + top.zgrid, self.nx, self.ny, self.nz)

    def nlevels(self):
        n = C.c_int64(0)
        N.call("wm_mg3d_nlevels", self._h, C.byref(n))
        return n.value

    def level_dims(self, level):
        nxyz = (C.c_int64 * 3)()
        dxyz = (C.c_double * 3)()
        N.call("wm_mg3d_level_dims", self._h, level, nxyz, dxyz)
        return tuple(nxyz), tuple(dxyz)


MultiGrid = MultiGrid3D
