"""find_mgparam: search for the SOR relaxation parameter (and pass counts)
that minimise the field-solve time, restating scripts/find_mgparam.py
(D. P. Grote's search): start from the current mgparam, step it by 0.05,
reverse and halve the step whenever the V-cycle count grows, stop when two
solves take the same number of V-cycles; then add down/up passes while the
solve time keeps falling.  The solver is any object with mgparam,
downpasses, uppasses, mgiters, mgmaxiters, mgtol, mgerror, phi, bounds and
solve(iwhich); MultiGrid3D.find_mgparam calls it on itself.
"""
import sys
import time

import numpy as np

_rng = np.random.default_rng(19950301)


def _log(*a):
    print(*a, file=sys.stderr)


def field_solve(phisave, solver):
    """One timed solve (find_mgparam.py:119-158): phi is zeroed over the
    points the solve updates, or restored from phisave."""
    phi = solver.phi
    if phisave is None:
        b = solver.bounds
        nx, ny, nz = solver.nx, solver.ny, solver.nz
        ixmin = 0 if b[0] > 0 else 1
        ixmax = nx if b[1] > 0 else nx - 1
        iymin = 0 if b[2] > 0 else 1
        iymax = ny if b[3] > 0 else ny - 1
        izmin = 0 if b[4] > 0 else 1
        izmax = nz if b[5] > 0 else nz - 1
        # phi is (nz+3, ny+3, nx+3) with one guard cell
        phi[izmin + 1:izmax + 2, iymin + 1:iymax + 2, ixmin + 1:ixmax + 2] = 0.
    else:
        phi[...] = phisave
    t0 = time.time()
    solver.solve(-1)
    return time.time() - t0


def _find_mgparam(phisave, solver):
    """find_mgparam.py:160-242."""
    icount = 0
    if solver.mgparam <= 0.:
        solver.mgparam = max(1., 2. + solver.mgparam)
    if solver.mgparam > 2.:
        solver.mgparam = max(1., 4. - solver.mgparam)
    fstime = field_solve(phisave, solver)
    mgparam_prev = solver.mgparam
    mgiters_prev = solver.mgiters
    sincr = .05
    solver.mgiters = 0
    solver.mgparam = solver.mgparam + sincr
    while mgiters_prev != solver.mgiters and icount < 20:
        _log("Best parameter so far = %f" % solver.mgparam)
        fstime = field_solve(phisave, solver)
        if solver.mgiters > mgiters_prev:
            sincr = -sincr / 2.
            solver.mgparam = mgparam_prev + sincr
        elif solver.mgiters < mgiters_prev:
            s = mgiters_prev
            mgiters_prev = solver.mgiters
            solver.mgiters = s
            mgparam_prev = solver.mgparam
            solver.mgparam = mgparam_prev + sincr
        if solver.mgparam <= 0.01 or 2. < solver.mgparam:
            solver.mgparam = 1.9 + _rng.random() * .05
            sincr = .01
        icount += 1
    if icount == 20:
        _log("Warning: maximum number of iterations reached.\n"
             "         The value of mgparam may not be optimal.")
    return fstime


def find_mgparam(solver, lsavephi=False, resetpasses=0):
    """find_mgparam.py:28-117: optimise mgparam and the pass counts,
    minimising the field-solve time; leaves the best values in the solver."""
    phisave = solver.phi.clone() if lsavephi else None
    if resetpasses:
        solver.downpasses = 1
        solver.uppasses = 1
    if solver.downpasses == 0:
        solver.downpasses = 1
    if solver.uppasses == 0:
        solver.uppasses = 1
    nexttime = field_solve(phisave, solver)
    prevtime = 2 * nexttime
    savemgtol = solver.mgtol
    if solver.mgiters == solver.mgmaxiters:
        solver.mgtol = 10 * solver.mgerror
        nexttime = field_solve(phisave, solver)
        prevtime = 2 * nexttime
    while nexttime < prevtime:
        prevparam = solver.mgparam
        prevtime = nexttime
        nexttime = _find_mgparam(phisave, solver)
        _log("Field solve time = ", nexttime)
        _log("f3d.mgparam = ", solver.mgparam)
        _log("f3d.downpasses = ", solver.downpasses)
        _log("f3d.uppasses = ", solver.uppasses)
        if nexttime < prevtime:
            solver.downpasses = solver.downpasses + 1
            solver.uppasses = solver.uppasses + 1
        else:
            solver.mgparam = prevparam
            solver.downpasses = solver.downpasses - 1
            solver.uppasses = solver.uppasses - 1
            if solver.downpasses == 0:
                solver.downpasses = 1
            if solver.uppasses == 0:
                solver.uppasses = 1
    if solver.mgiters == solver.mgmaxiters:
        _log("Notice: the maximum number of iterations has been reached, so the values "
             "above are unlikely to be optimal.")
    else:
        _log("The optimized values: time %g mgparam %g downpasses %d uppasses %d"
             % (prevtime, solver.mgparam, solver.downpasses, solver.uppasses))
    solver.mgtol = savemgtol
