"""GPU parity of the multigrid solver (libwarpmi through its C ABI) against
the FP64 oracle, and the reference's own multigrid known-answer test."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _solver(nx, ny, nz, lo, hi, bounds, **kw):
    from warp_amd import MultiGrid3D
    kw.setdefault("mgverbose", -1)
    return MultiGrid3D(nx=nx, ny=ny, nz=nz, xmmin=lo[0], xmmax=hi[0], ymmin=lo[1],
                       ymmax=hi[1], zmmin=lo[2], zmmax=hi[2], bounds=list(bounds), **kw)


@pytest.mark.parametrize("bounds", H.all_bounds())
def test_manufactured_kat(device, bounds):
    """MutltiGrid3D_test.py on the GPU: 24x22x18 on [-1,1]^3, mgtol 1e-12,
    error < 1e-10 and fewer than mgmaxiters V-cycles; and the same phi,
    bit for bit, and the same V-cycle count as the oracle."""
    nx, ny, nz = 24, 22, 18
    phie, rho, phi0, dx, dy, dz = H.manufactured(nx, ny, nz, bounds)
    s = _solver(nx, ny, nz, (-1, -1, -1), (1, 1, 1), bounds, mgtol=1e-12)
    # As the reference test's generate(): an initial (empty) loadrho leaves
    # the source finalized, so solve() does not fold rho at the faces again.
    s.loadrho()
    s.phi = torch.from_numpy(phi0).to(device)
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    phi = s.phi.cpu().numpy()
    po, ro = phi0.copy(), rho.copy()
    it, err = O.multigrid3dsolve(po, ro, dx, dy, dz, bounds=bounds, mgtol=1e-12)
    assert s.mgiters == it
    assert s.mgerror == err
    np.testing.assert_array_equal(phi, po)
    np.testing.assert_array_equal(s.rho.cpu().numpy(), ro)
    t = phie.copy()
    p = phi.copy()
    if 0 not in bounds:
        p -= p[1:-1, 1:-1, 1:-1].mean()
        t -= t[1:-1, 1:-1, 1:-1].mean()
    H.clearcorners(t)
    H.clearcorners(p)
    assert np.abs(p - t).max() < 1e-10
    assert s.mgiters < s.mgmaxiters


def _random_level_case(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    phi = rng.standard_normal((nz + 3, ny + 3, nx + 3))
    rho = rng.standard_normal((nz + 1, ny + 1, nx + 1))
    return phi, rho


@pytest.mark.parametrize("bounds", [(0,) * 6, (1, 1, 0, 1, 2, 2), (2, 2, 1, 0, 0, 1)])
@pytest.mark.parametrize("conductors", [False, True])
def test_kernels_match_oracle(device, bounds, conductors):
    """Each kernel alone on random data, level 0 and level 1, bitwise:
    sorhalfpass3d (both parities, with subgrid conductors), residual3d,
    restrict3d, expand3d."""
    nx, ny, nz = 32, 32, 64
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    s = _solver(nx, ny, nz, lo, hi, bounds)
    co = None
    if conductors:
        import warp_amd
        for g in H.pipe_and_annulus(warp_amd):
            s.installconductor(g)
        s.getconductorobject()
        co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
        O.checkconductors(co, nx, ny, nz, s.dx, s.dy, s.dz)
    from warp_amd import _native as N
    st = N.stream_of(device)
    for level in range(2):
        (lnx, lny, lnz), (ldx, ldy, ldz) = s.level_dims(level)
        phi, rho = _random_level_case(lnx, lny, lnz, 7 + level)
        for parity in (0, 1):
            for mgform in (1, 2):
                pg = torch.from_numpy(phi).to(device)
                rg = torch.from_numpy(rho).to(device)
                N.call("wm_mg3d_sorhalfpass", s._h, level, parity, pg.data_ptr(), rg.data_ptr(),
                       1.2, mgform, 1, 2, st)
                po = phi.copy()
                O.sorhalfpass3d(parity, po, rho, ldx, ldy, ldz, bounds, mgparam=1.2,
                                mgform=mgform, mglevel=level, conductors=co)
                np.testing.assert_array_equal(pg.cpu().numpy(), po)
        for mgform2init in (0, 1):
            pg = torch.from_numpy(phi).to(device)
            rg = torch.from_numpy(rho).to(device)
            res = torch.zeros_like(pg)
            N.call("wm_mg3d_residual", s._h, level, pg.data_ptr(), rg.data_ptr(), res.data_ptr(),
                   1.2, 1, mgform2init, 1, 2, st)
            ro = O.residual3d(phi, rho, ldx, ldy, ldz, bounds, mgparam=1.2, mgform=1,
                              mglevel=level, conductors=co, mgform2init=bool(mgform2init))
            np.testing.assert_array_equal(res.cpu().numpy(), ro)
        (cnx, cny, cnz), _ = s.level_dims(level + 1)
        ff = ((1. / ldx ** 2 + 1. / ldy ** 2 + 1. / ldz ** 2) /
              (1. / _[0] ** 2 + 1. / _[1] ** 2 + 1. / _[2] ** 2))
        resv = torch.from_numpy(ro).to(device)
        rc = torch.zeros((cnz + 1, cny + 1, cnx + 1), dtype=torch.float64, device=device)
        N.call("wm_mg3d_restrict", s._h, level, resv.data_ptr(), rc.data_ptr(), st)
        rco = O.restrict3d(ro, cnx, cny, cnz, ff, bounds)
        np.testing.assert_array_equal(rc.cpu().numpy(), rco)
        pc, _r = _random_level_case(cnx, cny, cnz, 11 + level)
        pf = torch.from_numpy(phi).to(device)
        N.call("wm_mg3d_expand", s._h, level, pf.data_ptr(),
               torch.from_numpy(pc).to(device).data_ptr(), st)
        pfo = phi.copy()
        O.expand3d(pfo, pc, bounds)
        np.testing.assert_array_equal(pf.cpu().numpy(), pfo)


@pytest.mark.parametrize("lwd,lzd", [(0, 0), (1, 0), (0, 1)])
@pytest.mark.parametrize("bounds", [(0,) * 6, (1,) * 6, (2,) * 6, (0, 1, 2, 2, 1, 0)])
def test_boundary_conditions(device, bounds, lwd, lzd):
    from warp_amd import _native as N
    nx, ny, nz = 10, 12, 14
    u = np.random.default_rng(3).standard_normal((nz + 3, ny + 3, nx + 3))
    ug = torch.from_numpy(u).to(device)
    b = (C.c_int64 * 6)(*bounds)
    N.call("wm_applyboundaryconditions3d", nx, ny, nz, ug.data_ptr(), b, lwd, lzd,
           N.stream_of(device))
    uo = u.copy()
    O.applyboundaryconditions3d(uo, bounds, lwd, lzd)
    np.testing.assert_array_equal(ug.cpu().numpy(), uo)


def _c1_like(device, nx, ny, nz, precision="fp64", mgform=1, seed=5):
    """Pipe + annulus with a Gaussian beam's rho; returns (solver, oracle
    phi, oracle mgiters)."""
    import warp_amd
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    s = _solver(nx, ny, nz, lo, hi, (0,) * 6, precision=precision, mgform=mgform)
    for g in H.pipe_and_annulus(warp_amd):
        s.installconductor(g)
    rng = np.random.default_rng(seed)
    from oracle import oracle as O2
    x = rng.normal(0, 0.004, 20000)
    y = rng.normal(0, 0.004, 20000)
    z = 0.04 + rng.normal(0, 0.015, 20000)
    k = (np.abs(x) < 0.0199) & (np.abs(y) < 0.0199) & (z > 0) & (z < 0.0799)
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O2.setrho3d(rho, x[k], y[k], z[k], 1.602176565e-19, 1e5, s.dx, s.dy, s.dz, lo[0], lo[1], lo[2])
    return s, rho


def test_conductor_solve_fp64_bitwise(device):
    """Full multigrid3dsolve with the pipe and annulus (subgrid conductors),
    64x64x128: same phi bit for bit and the same V-cycle count."""
    nx, ny, nz = 64, 64, 128
    s, rho = _c1_like(device, nx, ny, nz)
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    ro = rho.copy()
    it, err = O.multigrid3dsolve(po, ro, s.dx, s.dy, s.dz, conductors=co)
    assert s.mgiters == it and s.mgerror == err
    np.testing.assert_array_equal(s.phi.cpu().numpy(), po)
    # warm start from the converged phi: one more V-cycle
    s.solve()
    it2, err2 = O.multigrid3dsolve(po, ro, s.dx, s.dy, s.dz, conductors=co)
    assert s.mgiters == it2
    np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


@pytest.mark.parametrize("kind", ["m", "p", "mp", "dm"])
def test_neumann_subgrid_solve_bitwise(device, kind):
    """Hand-made Neumann subgrid lists (a sheet normal to x, dels in
    (-1, 0]; no conductor generator makes them): the solve runs the Neumann
    branches of the record kernels and the ghost mirror (condbndymgint).
    Its phi must be the closed form of
    test_oracle.py::test_neumann_sheet_closed_form (same case: the field
    g + h z, constant across the sheet) and agree with the oracle's to
    rounding, with the V-cycle count within one.  Unlike the Dirichlet
    paths this one is not bit for bit: phi differs from the oracle's in the
    last bits (mgerror 7.4358e-12 against 7.4216e-12 for "m", same count);
    the cause is not identified."""
    from warp_amd import _native as N
    n3, d3 = (12, 8, 16), (0.1, 0.12, 0.08)
    g, h, dl = 2.0, 25.0, 0.3
    vf = 1. - dl + dl * dl if kind == "dm" else 1.
    phi, _ = H.linear_z_case(n3, d3, g, h, seed=3)
    s = _solver(*n3, (0., 0., 0.), tuple(n * d for n, d in zip(n3, d3)), [0] * 6,
                mgtol=1e-11, mgmaxiters=200)
    s.getconductorobject()
    d = H.neumann_lists(n3, 5, kind, dl, lambda ix, iy, iz: (g + h * iz * d3[2]) * vf,
                        levels=s.nlevels())
    view = H.lists_struct(d)
    N.call("wm_mg3d_set_conductors", s._h, C.addressof(view))
    rho = np.zeros((n3[2] + 1, n3[1] + 1, n3[0] + 1))
    s.phi = torch.from_numpy(phi.copy()).to(device)
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    po = phi.copy()
    it, err = O.multigrid3dsolve(po, rho.copy(), *d3, conductors=H.oracle_conductors(d),
                                 mgtol=1e-11, mgmaxiters=200)
    assert it < 200 and s.mgerror <= 1e-11
    assert abs(s.mgiters - it) <= 1
    pg = s.phi.cpu().numpy()
    assert np.abs(pg - po).max() <= 1e-9
    _, exact = H.linear_z_case(n3, d3, g, h, seed=3)
    inner = (slice(1, n3[2] + 2), slice(1, n3[1] + 2), slice(1, n3[0] + 2))
    assert np.abs(pg[inner] - exact[inner]).max() < 1e-8


@pytest.mark.parametrize("down,up", [(2, 1), (1, 2), (2, 2), (0, 1), (1, 0)])
def test_conductor_solve_pass_counts_bitwise(device, down, up):
    """downpasses/uppasses other than 1 with the pipe and annulus, cold and
    warm solves: same phi, mgiters and mgerror bit for bit (the convergence
    test rides on the last up pass's sweeps after the first V-cycle, and is
    a separate launch when there is no up pass)."""
    nx, ny, nz = 32, 32, 64
    s, rho = _c1_like(device, nx, ny, nz)
    s.downpasses, s.uppasses = down, up
    s.rho = torch.from_numpy(rho).to(device)
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    ro = rho.copy()
    for k in range(2):
        if k == 1:
            # a changed source for the warm solve
            ro *= 1.25
            s.rho = torch.from_numpy(ro).to(device)
        s.solve()
        it, err = O.multigrid3dsolve(po, ro, s.dx, s.dy, s.dz, conductors=co,
                                     downpasses=down, uppasses=up)
        assert s.mgiters == it and s.mgerror == err, (k, s.mgiters, it, s.mgerror, err)
        np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


def test_solve_starts_from_a_changed_phi(device):
    """Consecutive solves with the pipe and annulus: after an untouched phi
    the solver starts from its own copy (wm_mg3d_reuse_phi); after the
    caller changed phi in place through a view, or through the setter, it
    starts from the changed phi.  Every solve bitwise equal to the oracle
    run on the same sequence."""
    nx, ny, nz = 32, 32, 64
    s, rho = _c1_like(device, nx, ny, nz)
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    ro = rho.copy()

    def both(tag, reused):
        s.rho = torch.from_numpy(ro).to(device)
        s.solve()
        it, err = O.multigrid3dsolve(po, ro, s.dx, s.dy, s.dz, conductors=co)
        assert (s.mgiters, s.mgerror) == (it, err), tag
        np.testing.assert_array_equal(s.phi.cpu().numpy(), po, err_msg=tag)
        assert s.last_solve_reused_phi() == reused, tag

    both("cold", False)
    ro *= 1.1
    both("untouched phi", True)
    s.phi[10:20, 5:9, 7:30] += 3.0    # in place, through a view of the tensor
    po[10:20, 5:9, 7:30] += 3.0
    ro *= 0.95
    both("changed through a view", False)
    po *= 0.5
    s.phi = torch.from_numpy(po).to(device)
    both("changed through the setter", False)
    both("untouched again", True)


def test_phi_reuse_across_the_two_buffer_form(device):
    """With one down pass or more, the V-cycles after the first keep their
    starting phi by exchanging the solver's two finest-level buffers, which
    leaves the save buffer stale at the end of the solve.  A following
    solve without down passes (which copies into save instead) must not
    start from that state; the one after it may again.  Every solve bitwise
    equal to the oracle run on the same sequence."""
    nx, ny, nz = 32, 32, 64
    s, rho = _c1_like(device, nx, ny, nz)
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    ro = rho.copy()

    def both(tag, down, reused):
        s.downpasses = down
        s.rho = torch.from_numpy(ro).to(device)
        s.solve()
        it, err = O.multigrid3dsolve(po, ro, s.dx, s.dy, s.dz, conductors=co, downpasses=down)
        assert (s.mgiters, s.mgerror) == (it, err), tag
        np.testing.assert_array_equal(s.phi.cpu().numpy(), po, err_msg=tag)
        assert s.last_solve_reused_phi() == reused, tag

    # with WARPMI_SAVE_SWAP=0 no solve exchanges the buffers, and every
    # untouched phi is reused
    swap = os.environ.get("WARPMI_SAVE_SWAP", "1").strip() != "0"
    both("cold", 1, False)
    for k, (down, reused) in enumerate([(1, True), (0, not swap), (0, True), (1, True),
                                        (2, True)]):
        ro *= 1.05
        both("solve %d, down %d" % (k, down), down, reused)


def test_level_entry_points_end_phi_reuse(device):
    """A per-level entry point between two solves (here the residual of
    another phi, which runs through the solver's level-0 buffers) must not
    let the next solve start from the stale copy: it converts the caller's
    phi again and stays bitwise equal to the oracle."""
    nx, ny, nz = 32, 32, 64
    s, rho = _c1_like(device, nx, ny, nz)
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    ro = rho.copy()
    s.rho = torch.from_numpy(ro).to(device)
    s.solve()
    O.multigrid3dsolve(po, ro.copy(), s.dx, s.dy, s.dz, conductors=co)
    keep = s.phi.clone()
    # the residual of a different phi, through the level-0 buffers
    s.phi = keep * 0.5
    s.getresidual()
    s.phi = keep          # phi now holds the last solve's result again ...
    s._phi_token = (s._phi._version, s._phi.data_ptr())   # ... and says so
    ro *= 1.1
    s.rho = torch.from_numpy(ro).to(device)
    s.solve()
    it, err = O.multigrid3dsolve(po, ro, s.dx, s.dy, s.dz, conductors=co)
    assert not s.last_solve_reused_phi()
    assert (s.mgiters, s.mgerror) == (it, err)
    np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


def test_conductor_solve_mgform2_bitwise(device):
    nx, ny, nz = 32, 32, 64
    s, rho = _c1_like(device, nx, ny, nz, mgform=2)
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    it, err = O.multigrid3dsolve(po, rho.copy(), s.dx, s.dy, s.dz, conductors=co, mgform=2)
    assert s.mgiters == it
    np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_conductor_solve_defect_correction(device, precision):
    """FP32 V-cycles (or BF16 on the finest level and FP32 below) in an
    FP64 defect-correction loop still reach mgtol and agree with the FP64
    oracle to the tolerance the convergence test implies."""
    nx, ny, nz = 64, 64, 128
    s, rho = _c1_like(device, nx, ny, nz, precision=precision)
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    assert s.mgerror <= s.mgtol
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    O.multigrid3dsolve(po, rho.copy(), s.dx, s.dy, s.dz, conductors=co)
    d = s.phi.cpu().numpy() - po
    # both are converged to max|dphi| <= 1e-6 V on a 10 kV problem
    assert np.abs(d).max() < 1e-4
    assert np.linalg.norm(d) / np.linalg.norm(po) < 1e-8


def test_nonconvergence_reported(device):
    """mgmaxiters = 1 returns mgerror > mgtol without raising
    (f3d_mgrid.F:460-463)."""
    nx, ny, nz = 32, 32, 64
    s, rho = _c1_like(device, nx, ny, nz)
    s.mgmaxiters = 1
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    assert s.mgiters == 1 and s.mgerror > s.mgtol


def test_quadrupole_channel_solve_bitwise(device):
    """The quadrupole channel's rods (ZCylinder, +-15 kV) inside the pipe:
    full FP64 solve, same phi bit for bit and the same V-cycle count as the
    oracle (32x32x64, Gaussian beam source)."""
    import warp_amd
    from warp_amd import configs
    nx, ny, nz = 32, 32, 64
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    s = _solver(nx, ny, nz, lo, hi, (0,) * 6)
    for g in configs.conductors("quads", warp_amd):
        s.installconductor(g)
    rng = np.random.default_rng(17)
    x, y = rng.normal(0, 0.003, 20000), rng.normal(0, 0.003, 20000)
    z = 0.04 + rng.normal(0, 0.015, 20000)
    k = (np.abs(x) < 0.0199) & (np.abs(y) < 0.0199) & (z > 0) & (z < 0.0799)
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O.setrho3d(rho, x[k], y[k], z[k], 1.602176565e-19, 5e5, s.dx, s.dy, s.dz, *lo)
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    co = O.install_conductors(configs.conductors("quads", O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    it, err = O.multigrid3dsolve(po, rho.copy(), s.dx, s.dy, s.dz, conductors=co)
    assert s.mgiters == it and s.mgerror == err
    np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


def test_cropped_conductor_solve_bitwise(device):
    """installconductor with a region (the annulus generated only for
    z in [3.1, 5] cm, the pipe everywhere): full FP64 solve bitwise equal to
    the oracle's with the same cropped lists."""
    import warp_amd
    nx, ny, nz = 32, 32, 64
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    s = _solver(nx, ny, nz, lo, hi, (0,) * 6)
    pipe, ann = H.pipe_and_annulus(warp_amd)
    s.installconductor(pipe)
    s.installconductor(ann, zmin=0.031, zmax=0.05)
    rng = np.random.default_rng(41)
    x, y = rng.normal(0, 0.004, 20000), rng.normal(0, 0.004, 20000)
    z = 0.04 + rng.normal(0, 0.015, 20000)
    k = (np.abs(x) < 0.0199) & (np.abs(y) < 0.0199) & (z > 0) & (z < 0.0799)
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O.setrho3d(rho, x[k], y[k], z[k], 1.602176565e-19, 1e5, s.dx, s.dy, s.dz, *lo)
    s.rho = torch.from_numpy(rho).to(device)
    s.solve()
    opipe, oann = H.pipe_and_annulus(O)
    co = O.install_conductors([opipe], nx, ny, nz, *H.DOMAIN)
    co = O.install_conductors([oann], nx, ny, nz, *H.DOMAIN, conductors=co,
                              crop=(None, None, None, None, 0.031, 0.05))
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    it, err = O.multigrid3dsolve(po, rho.copy(), s.dx, s.dy, s.dz, conductors=co)
    assert s.mgiters == it and s.mgerror == err
    np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


@pytest.mark.parametrize("case", ["pipe_annulus", "quads_periodic", "neumann_periodic"])
def test_vcycle_tail_same_bits(device, case, monkeypatch):
    """The coarse end of the V-cycle as one launch (levels in LDS) against
    the level-by-level launches: the same phi bit for bit and the same
    V-cycle counts, cold and warm; the single launch did run."""
    import warp_amd
    from warp_amd import configs
    nx, ny, nz = 64, 64, 128
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    bounds = {"pipe_annulus": (0,) * 6, "quads_periodic": (0, 0, 0, 0, 2, 2),
              "neumann_periodic": (1, 1, 0, 1, 2, 2)}[case]
    conds = (H.pipe_and_annulus(warp_amd) if case == "pipe_annulus" else
             configs.conductors("quads", warp_amd) if case == "quads_periodic" else [])
    rng = np.random.default_rng(31)
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    x, y = rng.normal(0, 0.004, 30000), rng.normal(0, 0.004, 30000)
    z = 0.04 + rng.normal(0, 0.015, 30000)
    k = (np.abs(x) < 0.0199) & (np.abs(y) < 0.0199) & (z > 0) & (z < 0.0799)
    out = []
    for budget in ("0", None):
        if budget is None:
            monkeypatch.delenv("WARPMI_TAIL_LDS", raising=False)
        else:
            monkeypatch.setenv("WARPMI_TAIL_LDS", budget)
        s = _solver(nx, ny, nz, lo, hi, bounds)
        for g in conds:
            s.installconductor(g)
        if not out:
            O.setrho3d(rho, x[k], y[k], z[k], 1.602176565e-19, 5e5, s.dx, s.dy, s.dz, *lo)
        res = []
        for f in (1.0, 1.1):
            s.rho = torch.from_numpy(rho * f).to(device)
            s.solve()
            res.append((s.mgiters, s.mgerror, s.phi.cpu().numpy().copy(), s.tail_start_level()))
        out.append(res)
    for a, b in zip(*out):
        assert a[3] == -1 and b[3] >= 1, (a[3], b[3])
        assert (a[0], a[1]) == (b[0], b[1])
        np.testing.assert_array_equal(a[2], b[2])


@pytest.mark.parametrize("bounds", [(0,) * 6, (1, 1, 0, 1, 2, 2), (2, 2, 1, 0, 0, 1)])
def test_residual_wide_rows_match_oracle(device, bounds):
    """residual3d on a level with padded rows (nx = 128: 80 elements per
    half row), where the FP64 residual takes two points of each colour per
    thread with 16-byte loads: bitwise equal to the oracle, with the pipe
    and annulus, both initial-residual forms."""
    import warp_amd
    nx, ny, nz = 128, 16, 24
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    s = _solver(nx, ny, nz, lo, hi, bounds)
    for g in H.pipe_and_annulus(warp_amd):
        s.installconductor(g)
    s.getconductorobject()
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    O.checkconductors(co, nx, ny, nz, s.dx, s.dy, s.dz)
    from warp_amd import _native as N
    st = N.stream_of(device)
    phi, rho = _random_level_case(nx, ny, nz, 41)
    for mgform2init in (0, 1):
        pg = torch.from_numpy(phi).to(device)
        rg = torch.from_numpy(rho).to(device)
        res = torch.zeros_like(pg)
        N.call("wm_mg3d_residual", s._h, 0, pg.data_ptr(), rg.data_ptr(), res.data_ptr(),
               1.2, 1, mgform2init, 1, 2, st)
        ro = O.residual3d(phi, rho, s.dx, s.dy, s.dz, bounds, mgparam=1.2, mgform=1,
                          mglevel=0, conductors=co, mgform2init=bool(mgform2init))
        np.testing.assert_array_equal(res.cpu().numpy(), ro)
