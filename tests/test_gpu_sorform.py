"""The fused SOR pass (both colours of sorpass3d, f3d_mgrid.F:1949-2093, in
one launch reading one red-black phi buffer and writing the other) against
the two half sweeps and the FP64 oracle: the same phi bit for bit and the
same V-cycle counts, cold and warm, with and without conductors, at the
reference's pass counts and at others, and in the defect-correction
precisions (where the fused pass runs on the FP32/BF16 levels)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _solver(n, sorform, conductors, **kw):
    import warp_amd
    from warp_amd import MultiGrid3D
    nx, ny, nz = n
    lo, hi = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.08)
    s = MultiGrid3D(nx=nx, ny=ny, nz=nz, xmmin=lo[0], xmmax=hi[0], ymmin=lo[1], ymmax=hi[1],
                    zmmin=lo[2], zmmax=hi[2], bounds=[0] * 6, mgverbose=-1, sorform=sorform,
                    **kw)
    if conductors:
        for g in H.pipe_and_annulus(warp_amd):
            s.installconductor(g)
    return s


def _rho(n, seed):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    x, y = rng.normal(0, 0.004, 40000), rng.normal(0, 0.004, 40000)
    z = 0.04 + rng.normal(0, 0.015, 40000)
    k = (np.abs(x) < 0.0199) & (np.abs(y) < 0.0199) & (z > 0) & (z < 0.0799)
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O.setrho3d(rho, x[k], y[k], z[k], 1.602176565e-19, 1e5, 0.04 / nx, 0.04 / ny, 0.08 / nz,
               -0.02, -0.02, 0.0)
    return rho


@pytest.mark.parametrize("n", [(32, 32, 64), (64, 64, 128), (48, 40, 96)])
@pytest.mark.parametrize("conductors", [False, True])
def test_fused_pass_equals_oracle(device, n, conductors):
    """Cold solve, then two warm solves on changed sources (the second
    reusing the solver's own phi): bitwise phi, mgiters, mgerror."""
    nx, ny, nz = n
    s = _solver(n, 1, conductors)
    rho = _rho(n, 21)
    co = (O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
          if conductors else None)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    for k, f in enumerate((1.0, 1.2, 0.9)):
        ro = rho * f
        s.rho = torch.from_numpy(ro).to(device)
        s.solve()
        it, err = O.multigrid3dsolve(po, ro.copy(), s.dx, s.dy, s.dz, conductors=co)
        assert (s.mgiters, s.mgerror) == (it, err), (k, s.mgiters, it, s.mgerror, err)
        np.testing.assert_array_equal(s.phi.cpu().numpy(), po, err_msg=str(k))


@pytest.mark.parametrize("down,up", [(1, 1), (2, 2), (1, 3), (2, 1)])
def test_fused_pass_pass_counts(device, down, up):
    """Pass counts other than the reference's: with an even total the
    fused pass runs, with an odd one the half sweeps do; both bitwise."""
    n = (32, 32, 64)
    nx, ny, nz = n
    s = _solver(n, 1, True, downpasses=down, uppasses=up)
    rho = _rho(n, 22)
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    for k, f in enumerate((1.0, 1.3)):
        ro = rho * f
        s.rho = torch.from_numpy(ro).to(device)
        s.solve()
        it, err = O.multigrid3dsolve(po, ro.copy(), s.dx, s.dy, s.dz, conductors=co,
                                     downpasses=down, uppasses=up)
        assert (s.mgiters, s.mgerror) == (it, err), (k, down, up)
        np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_pass_defect_correction_same_bits(device, precision):
    """Defect correction with the fused pass on its low-precision levels:
    the same phi and V-cycles as with the half sweeps."""
    n = (64, 64, 128)
    rho = _rho(n, 23)
    out = []
    for form in (0, 1):
        s = _solver(n, form, True, precision=precision)
        its = []
        for f in (1.0, 1.1):
            s.rho = torch.from_numpy(rho * f).to(device)
            s.solve()
            its.append((s.mgiters, s.mgerror))
        out.append((its, s.phi.cpu().numpy()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("conductors", [False, True])
def test_fused_finest_level_only(device, conductors):
    """Form 2 (the fused pass on the finest level, half sweeps below):
    bitwise equal to the oracle, cold and warm."""
    n = (64, 64, 128)
    nx, ny, nz = n
    s = _solver(n, 2, conductors)
    rho = _rho(n, 25)
    co = (O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
          if conductors else None)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    for k, f in enumerate((1.0, 0.8)):
        ro = rho * f
        s.rho = torch.from_numpy(ro).to(device)
        s.solve()
        it, err = O.multigrid3dsolve(po, ro.copy(), s.dx, s.dy, s.dz, conductors=co)
        assert (s.mgiters, s.mgerror) == (it, err), k
        np.testing.assert_array_equal(s.phi.cpu().numpy(), po)


def test_switching_forms_between_solves(device):
    """wm_mg3d_set_sor_form between solves of one solver: every solve
    still equals the oracle bit for bit."""
    from warp_amd import _native as N
    n = (32, 32, 64)
    nx, ny, nz = n
    s = _solver(n, 0, True)
    rho = _rho(n, 24)
    co = O.install_conductors(H.pipe_and_annulus(O), nx, ny, nz, *H.DOMAIN)
    po = np.zeros((nz + 3, ny + 3, nx + 3))
    for k, form in enumerate((0, 1, 1, 0, 2, 1)):
        N.call("wm_mg3d_set_sor_form", s._h, form)
        ro = rho * (1. + 0.1 * k)
        s.rho = torch.from_numpy(ro).to(device)
        s.solve()
        it, err = O.multigrid3dsolve(po, ro.copy(), s.dx, s.dy, s.dz, conductors=co)
        assert (s.mgiters, s.mgerror) == (it, err), (k, form)
        np.testing.assert_array_equal(s.phi.cpu().numpy(), po)
