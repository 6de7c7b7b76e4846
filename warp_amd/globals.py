"""Package-level inputs mirroring the Forthon packages the reference solver
reads its defaults from (pywarp90/top.v, w3d.v, f3d.v): ``top``, ``w3d``
and ``f3d``.  Only the variables the electrostatic PIC path uses exist.
"""
import math


class _Package:
    def __init__(self, **kw):
        self.__dict__.update(kw)

    def __repr__(self):
        return "%s(%s)" % (type(self).__name__, ", ".join(
            "%s=%r" % kv for kv in sorted(self.__dict__.items()) if not kv[0].startswith("_")))


# Physical constants (pywarp90/top.v:172-181; scripts/warp.py:198-202).
pi = 3.14159265358979323
amu = 1.660538921e-27
clight = 2.99792458e+8
echarge = 1.602176565e-19
emass = 9.10938291e-31
mu0 = 4 * pi / 10000000
eps0 = 1 / (mu0 * clight * clight)
largepos = 1.0e+36
dirichlet, neumann, periodic = 0, 1, 2


class Top(_Package):
    pass


class W3D(_Package):
    pass


class F3D(_Package):
    pass


# top: time step and particle data (top.v); fstype=12 once a solver is
# registered (fieldsolver.py:181).
top = Top(dt=0., time=0., it=0, fstype=-1, pgroup=None, zbeam=0., zgrid=0.,
          depos="direct1", efetch=1, lresetparticlee=True, steptime=0.,
          fstime=0., lrtime=0.,
          # steps between particlesortxyzwithcopy calls in step(); 0 = never
          particlesortinterval=0,
          # clearpart's fill method for lost particles (top.v:1260): 1 fills
          # holes with particles from the end, 2 keeps the order, 0 leaves them
          clearlostpart=1,
          # particle boundaries (top.v:1187-1195): at iz = 0, at iz = nz, in x
          # and y; absorb = 0, reflect = 1, periodic = 2 (top.v:75-78)
          pbound0=2, pboundnz=2, pboundxy=0,
          # step controls of w3dexe/step3d (top.v): "special" steps end with
          # the synchv half kick and the beam moments
          nt=20000000, tstop=1.0e36, zstop=1.0e36, itmomnts=[0] * 50, nhist=5, itdump=500,
          allspecl=False, lspecial=False, nspecial=0,
          # beam moments of the last special step, and the moment history
          # (every nhist steps): dicts keyed by top's variable names
          moments={}, hist=[],
          # Z_arrays mesh of the line charge (top.v:1615-1629; set from the
          # field mesh by w3d.F:337-341 when left zero) and the field
          # diagnostics of special steps: line charge [C/m] and
          # electrostatic energy [J] (top.v:1898)
          nzzarr=0, zzmin=0., zzmax=0., dzz=0., linechg=None, ese=0.)

# w3d: mesh and boundary inputs (w3d.v), plus the FieldSolveAPI globals
# (w3d.v:432-458) through which fetche receives particle positions.
w3d = W3D(nx=2, ny=2, nz=2, xmmin=0., xmmax=1., ymmin=0., ymmax=1., zmmin=0., zmmax=1.,
          nxguardphi=1, nyguardphi=1, nzguardphi=1, nxguardrho=0, nyguardrho=0,
          nzguardrho=0, nxguarde=0, nyguarde=0, nzguarde=0,
          bound0=dirichlet, boundnz=dirichlet, boundxy=dirichlet,
          l2symtry=False, l4symtry=False,
          # field diagnostics switches (w3d.v:70-74)
          lgetese3d=True, lgtlchg3d=True,
          # FieldSolveAPI
          npfsapi=0, xfsapi=None, yfsapi=None, zfsapi=None, exfsapi=None,
          eyfsapi=None, ezfsapi=None, phifsapi=None, pgroupfsapi=None,
          ipminfsapi=1, jsfsapi=-1, ndtsfsapi=1, lzerorhofsapi=True,
          lfinalizerhofsapi=True, js1fsapi=-1, js2fsapi=-1)

# f3d: multigrid controls with their defaults (f3d.v:169-231).
f3d = F3D(gridmode=1, mgparam=1.2, mgmaxiters=100, mgmaxlevels=101, mgtol=1.e-6,
          mgform=1, downpasses=1, uppasses=1, mgverbose=1, mgntverbose=1,
          lcndbndy=True, icndbndy=2, laddconductor=False, lprecalccoeffs=False,
          mgiters=0, mgerror=0.,
          # E near conductors from one-sided differences (f3d.v:152-153)
          lcorrectede=False)

assert math.isclose(eps0, 8.854187817620389e-12, rel_tol=1e-15)
