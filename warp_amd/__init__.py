"""warp_amd: MI355X (gfx950) implementation of WARP's 3-D electrostatic PIC
hot path, as a drop-in registered field solver.

    from warp_amd import MultiGrid3D, registersolver, ZCylinderOut
    solver = MultiGrid3D(nx=128, ny=128, nz=256, xmmin=-0.02, ...)
    registersolver(solver)
    solver.installconductor(ZCylinderOut(0.018, 0.1, zcent=0.04))
    solver.loadrho(); solver.solve(); solver.fetche()

The native library is libwarpmi.so (include/warpmi.h), built in-tree by
``python -m warp_amd.build``.
"""
from .globals import top, w3d, f3d, eps0, amu, echarge, clight, pi, largepos
from .fieldsolver import (registersolver, unregistersolver, getregisteredsolver,
                          getregisteredsolvers, loadrhoregistered, fieldsolregistered,
                          fetcheregistered, fetchphiregistered, fieldsolve, FieldSolver)
from .generateconductors import (Assembly, AssemblyPlus, ZCylinder, ZCylinderOut,
                                 ZSrfrvIn, ZAnnulus, QuadRods, Grid, ConductorData,
                                 installconductor, getmglevels)


def __getattr__(name):
    # Modules that need torch are imported on first use so that the
    # conductor set-up stays usable without it.
    if name in ("MultiGrid3D", "MultiGrid"):
        from . import multigrid
        return getattr(multigrid, name)
    if name in ("ParticleGroup", "load_gaussian", "load_kv_ellipse"):
        from . import particles
        return getattr(particles, name)
    if name in ("step", "generate"):
        from . import stepper
        return getattr(stepper, name)
    raise AttributeError(name)
