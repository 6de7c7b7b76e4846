"""Registered-solver plugin API, mirroring scripts/fieldsolver.py.

``registersolver`` appends a solver to the list the time step calls
(fieldsolver.py:173-229): ``loadrhoregistered``, ``fieldsolregistered``,
``fetcheregistered`` and ``fetchphiregistered`` call the method of every
registered solver in order.  :class:`FieldSolver` holds the mesh inputs
with the package defaults (processdefaultsfrompackage,
fieldsolver.py:451-456) and raises on unknown keyword arguments.
"""
import copy

from .globals import top, w3d

_registeredfieldsolvers = []


def registersolver(solver):
    """fieldsolver.py:173-182: the solver must define loadrho, solve and
    fetche (fetchphi when potentials are fetched)."""
    _registeredfieldsolvers.append(solver)
    top.fstype = 12


def _checkfstypeforinvalidvalue():
    if len(_registeredfieldsolvers) > 0 and top.fstype != 12:
        raise RuntimeError('top.fstype should not be changed when a field solver has been registered')


def getregisteredsolver(i=0):
    if len(_registeredfieldsolvers) == 0:
        return None
    _checkfstypeforinvalidvalue()
    return _registeredfieldsolvers[i]


def getregisteredsolvers():
    _checkfstypeforinvalidvalue()
    return copy.copy(_registeredfieldsolvers)


def unregistersolver(solver=None, i=None):
    _checkfstypeforinvalidvalue()
    if i is not None:
        del _registeredfieldsolvers[i]
    else:
        if solver is None:
            solver = _registeredfieldsolvers[0]
        _registeredfieldsolvers.remove(solver)
    if not _registeredfieldsolvers:
        top.fstype = -1


def _each(method):
    assert len(_registeredfieldsolvers) > 0, "No solver has been registered"
    for f in _registeredfieldsolvers:
        getattr(f, method)()


def loadrhoregistered():
    _each("loadrho")


def fieldsolregistered():
    _each("solve")


def fetcheregistered():
    _each("fetche")


def fetchphiregistered():
    _each("fetchphi")


# The diagnostics step3d calls through the registered solvers
# (fieldsolver.py:234-257; w3d.F:987-1002): rhodia every step, the others
# on special steps.
def rhodiaregistered():
    _each("rhodia")


def gtlchgregistered():
    _each("gtlchg")


def srhoaxregistered():
    _each("srhoax")


def geteseregistered():
    _each("getese")


def sphiaxregistered():
    _each("sphiax")


def sezaxregistered():
    _each("sezax")


def fieldsolve(iwhich=0):
    """scripts/fieldsolver.py:49-89 for a registered solver."""
    for f in getregisteredsolvers():
        f.solve(iwhich)


class FieldSolver:
    """Base of registered field solvers (fieldsolver.py:378-456)."""

    __w3dinputs__ = ['nx', 'ny', 'nz', 'xmmin', 'xmmax', 'ymmin', 'ymmax', 'zmmin', 'zmmax',
                     'nxguardphi', 'nyguardphi', 'nzguardphi',
                     'nxguardrho', 'nyguardrho', 'nzguardrho',
                     'nxguarde', 'nyguarde', 'nzguarde',
                     'bound0', 'boundnz', 'boundxy', 'l2symtry', 'l4symtry']
    __flaginputs__ = {'ldosolve': 1, 'l_internal_dosolve': 1, 'debug': 0}

    def __init__(self, kwdict=None, **kw):
        # As in the reference (fieldsolver.py:432-437), subclasses pass their
        # keyword dict as kwdict so that consumed entries are removed from it.
        if kwdict is not None:
            kwdict.update(kw)
            kw = kwdict
        self.processdefaultsfrompackage(FieldSolver.__w3dinputs__, w3d, kw)
        self.processdefaultsfromdict(FieldSolver.__flaginputs__, kw)
        if self.l2symtry or self.l4symtry:
            raise NotImplementedError("transverse symmetries are not supported")
        for g in ('nxguardphi', 'nyguardphi', 'nzguardphi'):
            if getattr(self, g) != 1:
                raise NotImplementedError("phi must have exactly one guard cell")
        for g in ('nxguardrho', 'nyguardrho', 'nzguardrho', 'nxguarde', 'nyguarde', 'nzguarde'):
            if getattr(self, g) != 0:
                raise NotImplementedError("rho and E arrays have no guard cells")
        self.setupbounds(kw)
        self.dx = (self.xmmax - self.xmmin) / self.nx
        self.dy = (self.ymmax - self.ymmin) / self.ny
        self.dz = (self.zmmax - self.zmmin) / self.nz
        self.nxlocal, self.nylocal, self.nzlocal = self.nx, self.ny, self.nz
        self.xmminlocal, self.ymminlocal, self.zmminlocal = self.xmmin, self.ymmin, self.zmmin
        # particle-side mesh equals the field mesh on one GPU
        self.nxp, self.nyp, self.nzp = self.nx, self.ny, self.nz
        self.xmminp, self.ymminp, self.zmminp = self.xmmin, self.ymmin, self.zmmin
        self.xmmaxp, self.ymmaxp, self.zmmaxp = self.xmmax, self.ymmax, self.zmmax
        self.sourcepfinalized = False

    def processdefaultsfrompackage(self, defaults, package, kw):
        for name in defaults:
            if name not in self.__dict__:
                self.__dict__[name] = kw.get(name, getattr(package, name))
            if name in kw:
                del kw[name]

    def processdefaultsfromdict(self, d, kw):
        for name, defvalue in d.items():
            if name not in self.__dict__:
                self.__dict__[name] = kw.get(name, defvalue)
            if name in kw:
                del kw[name]

    def setupbounds(self, kw):
        """fieldsolver.py setupbounds: bounds from bound0, boundnz, boundxy
        unless given explicitly."""
        if 'bounds' in kw:
            self.bounds = [int(b) for b in kw.pop('bounds')]
        else:
            self.bounds = [self.boundxy, self.boundxy, self.boundxy, self.boundxy,
                           self.bound0, self.boundnz]
        assert 0 <= min(self.bounds) and max(self.bounds) <= 2, \
            "The boundary conditions have an incorrect value. They must be one of dirichlet, neumann or periodic."
        for a in range(3):
            if self.bounds[2 * a] == 2 or self.bounds[2 * a + 1] == 2:
                self.bounds[2 * a] = self.bounds[2 * a + 1] = 2

    # The per-step interface (fieldsolver.py:379-390).
    def loadrho(self, lzero=True, lfinalize_rho=True, pgroups=None, **kw):
        raise NotImplementedError

    def solve(self, iwhich=0):
        raise NotImplementedError

    def fetche(self, *args, **kw):
        raise NotImplementedError

    def fetchphi(self, *args, **kw):
        raise NotImplementedError

    def installconductor(self, conductor, xmin=None, xmax=None, ymin=None, ymax=None,
                         zmin=None, zmax=None, dfill=None):
        raise NotImplementedError

    # Diagnostics a solver may provide (fieldsolver.py:1053-1064): none by
    # default.
    def rhodia(self):
        pass

    def gtlchg(self):
        pass

    def srhoax(self):
        pass

    def getese(self):
        pass

    def sphiax(self):
        pass

    def sezax(self):
        pass
