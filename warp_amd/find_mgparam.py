"""find_mgparam: tune the SOR relaxation parameter mgparam and the number
of down/up passes of a multigrid solver for the shortest field solve.

The interface is the reference's (scripts/find_mgparam.py:28-117, called by
MultiGrid3D.find_mgparam, multigrid.py:827-831): the solver is any object
with mgparam, downpasses, uppasses, mgiters, mgmaxiters, mgtol, mgerror,
phi, bounds, nx, ny, nz and solve(iwhich); with lsavephi every trial solve
starts from the phi the solver holds on entry, otherwise from zero on the
points the solve updates; with resetpasses the pass counts start at 1; a
tolerance the solve cannot reach within mgmaxiters is loosened to ten times
the error reached while searching, then restored.  The best values are left
in the solver.

The search is this module's own: for each pass count (1, 2, ... while the
solve keeps getting faster) mgparam is scanned on a coarse grid over
[1.0, 1.95] and then refined around the best point; a candidate is better
when it needs fewer V-cycles, or as many in less time.  It costs a bounded
number of solves (about 30 per pass count) and uses no random numbers.
"""
import sys
import time

COARSE = [round(1.0 + 0.05 * k, 2) for k in range(20)]  # 1.00 .. 1.95
FINE_STEP = 0.01
FINE_HALF_WIDTH = 4   # fine points on each side of the coarse optimum
MAX_PASSES = 8


def _log(*a):
    print(*a, file=sys.stderr)


class _Trial:
    """One timed solve from the chosen starting phi."""

    def __init__(self, solver, lsavephi):
        self.solver = solver
        self.start = solver.phi.clone() if lsavephi else None

    def _reset_phi(self):
        s = self.solver
        if self.start is not None:
            s.phi[...] = self.start
            return
        # zero the points the solve updates: Dirichlet faces keep their values
        lo = [0 if b > 0 else 1 for b in s.bounds[0::2]]
        hi = [n if b > 0 else n - 1 for n, b in zip((s.nx, s.ny, s.nz), s.bounds[1::2])]
        # phi is (nz+3, ny+3, nx+3) with one guard cell on each side
        s.phi[lo[2] + 1:hi[2] + 2, lo[1] + 1:hi[1] + 2, lo[0] + 1:hi[0] + 2] = 0.

    def __call__(self, mgparam=None):
        s = self.solver
        if mgparam is not None:
            s.mgparam = mgparam
        self._reset_phi()
        t0 = time.perf_counter()
        s.solve(-1)
        return s.mgiters, time.perf_counter() - t0


def _better(a, b):
    """Fewer V-cycles wins; equal counts go to the faster solve."""
    return b is None or a[0] < b[0] or (a[0] == b[0] and a[1] < b[1])


def tune_mgparam(trial):
    """The mgparam with the fewest V-cycles (then the shortest time) for the
    solver's current pass counts: a coarse scan, then a fine one around
    its best point.  Returns (mgparam, (mgiters, seconds))."""
    best_p, best = None, None
    for p in COARSE:
        r = trial(p)
        if _better(r, best):
            best_p, best = p, r
    fine = [round(best_p + FINE_STEP * k, 2) for k in range(-FINE_HALF_WIDTH, FINE_HALF_WIDTH + 1)
            if k != 0]
    for p in fine:
        if not 0. < p < 2.:
            continue
        r = trial(p)
        if _better(r, best):
            best_p, best = p, r
    trial(best_p)  # leave the solver holding the best solve
    return best_p, best


def find_mgparam(solver, lsavephi=False, resetpasses=0):
    """Tunes solver.mgparam, then the pass counts (downpasses = uppasses),
    for the shortest solve time; leaves the best values in the solver."""
    trial = _Trial(solver, lsavephi)
    if resetpasses:
        solver.downpasses = solver.uppasses = 1
    solver.downpasses = max(1, solver.downpasses)
    solver.uppasses = max(1, solver.uppasses)
    savemgtol = solver.mgtol
    trial()
    if solver.mgiters >= solver.mgmaxiters:
        # the tolerance is out of reach: search at what the solve reaches
        solver.mgtol = 10 * solver.mgerror
    best = None  # (seconds, mgparam, down, up, mgiters)
    down, up = solver.downpasses, solver.uppasses
    while down <= MAX_PASSES:
        solver.downpasses, solver.uppasses = down, up
        p, (iters, secs) = tune_mgparam(trial)
        _log("find_mgparam: passes %d/%d mgparam %.2f -> %d V-cycles, %.3g s"
             % (down, up, p, iters, secs))
        if best is not None and secs >= best[0]:
            break
        best = (secs, p, down, up, iters)
        down, up = down + 1, up + 1
    secs, solver.mgparam, solver.downpasses, solver.uppasses, iters = best
    solver.mgtol = savemgtol
    if iters >= solver.mgmaxiters:
        _log("find_mgparam: the solve still reaches mgmaxiters; the values found are "
             "unlikely to be optimal")
    else:
        _log("find_mgparam: time %g s, mgparam %g, downpasses %d, uppasses %d"
             % (secs, solver.mgparam, solver.downpasses, solver.uppasses))
