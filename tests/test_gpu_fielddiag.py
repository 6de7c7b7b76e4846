"""GPU parity of the field diagnostics (line charge gtlchg3dfromrho and
electrostatic energy getese3dfromrhophi) through the C ABI against the FP64
oracle: bitwise, for every symmetry mode, ragged and degenerate meshes,
planes longer than one staging chunk, and the 256x256x512 mesh."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SYMS = [(False, False), (True, False), (False, True), (True, True)]


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def _gtlchg_gpu(device, rho, dx, dy, dz, zgrid, zmmin, zzmin, zbeam, dzz, nzzarr, l2, l4):
    from warp_amd import _native as N
    nz, ny, nx = (d - 1 for d in rho.shape)
    r = _t(rho, device)
    planes = torch.full((nz + 1,), float("nan"), dtype=torch.float64, device=device)
    lc = torch.full((nzzarr + 1,), float("nan"), dtype=torch.float64, device=device)
    N.call("wm_gtlchg3d", r.data_ptr(), nx, ny, nz, dx, dy, dz, zgrid, zmmin, zzmin, zbeam, dzz,
           nzzarr, int(l2), int(l4), planes.data_ptr(), lc.data_ptr(), N.stream_of(device))
    return lc.cpu().numpy(), planes.cpu().numpy()


def _getese_gpu(device, rho, phi, dx, dy, dz, l4, l2):
    from warp_amd import _native as N
    nz, ny, nx = (d - 1 for d in rho.shape)
    r, p = _t(rho, device), _t(phi, device)
    work = torch.empty(int(N.lib().wm_getese3d_worksize(nx, ny)), dtype=torch.uint8,
                       device=device)
    out = torch.full((1,), float("nan"), dtype=torch.float64, device=device)
    N.call("wm_getese3d", r.data_ptr(), p.data_ptr(), nx, ny, nz, dx, dy, dz, int(l4), int(l2),
           work.data_ptr(), out.data_ptr(), N.stream_of(device))
    return float(out.item())


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


@pytest.mark.parametrize("n3", [(9, 7, 12), (0, 0, 0), (4, 0, 3), (0, 5, 2), (63, 65, 5),
                                (40, 50, 1)])
@pytest.mark.parametrize("l2,l4", SYMS)
def test_gtlchg_bitwise(device, n3, l2, l4):
    """Plane charges and line charge equal the oracle's bits.  (63, 65)
    planes hold 4224 nodes: three staging chunks, the last one partial."""
    nx, ny, nz = n3
    rng = np.random.default_rng(nx * 1000 + ny * 10 + nz)
    rho = rng.standard_normal((nz + 1, ny + 1, nx + 1)) * 1e-9
    dx, dy, dz = 0.1, 0.2, 0.05
    for zbeam, nzzarr, dzz in ((0., nz, dz), (0.013, 2 * nz + 3, 0.6 * dz), (-0.4, 4, 2.3 * dz)):
        args = (rho, dx, dy, dz, 0.002, 0.1, 0.08, zbeam, dzz, nzzarr)
        lc, planes = _gtlchg_gpu(device, *args, l2, l4)
        lco, planeso = O.gtlchg3d(*args, l2symtry=l2, l4symtry=l4)
        assert _bits_equal(planes, planeso)
        assert _bits_equal(lc, lco)


@pytest.mark.parametrize("n3", [(9, 7, 12), (0, 0, 0), (4, 0, 3), (63, 65, 5), (17, 3, 40)])
@pytest.mark.parametrize("l2,l4", SYMS)
def test_getese_bitwise(device, n3, l2, l4):
    nx, ny, nz = n3
    rng = np.random.default_rng(nx * 1000 + ny * 10 + nz + 1)
    rho = rng.standard_normal((nz + 1, ny + 1, nx + 1))
    phi = rng.standard_normal((nz + 3, ny + 3, nx + 3))
    e = _getese_gpu(device, rho, phi, 0.1, 0.2, 0.05, l4, l2)
    eo = O.getese3d(rho, phi, 0.1, 0.2, 0.05, l4symtry=l4, l2symtry=l2)
    assert _bits_equal(e, eo), (e, eo)


def test_full_mesh_bitwise_and_closed_form(device):
    """256x256x512: bitwise against the oracle, and the sums against
    numpy's (a different order) to 1e-12."""
    nx, ny, nz = 256, 256, 512
    dx, dy, dz = 0.04 / nx, 0.04 / ny, 0.08 / nz
    rng = np.random.default_rng(7)
    rho = rng.uniform(0., 1e-6, (nz + 1, ny + 1, nx + 1))
    phi = rng.uniform(0., 1e4, (nz + 3, ny + 3, nx + 3))
    lc, planes = _gtlchg_gpu(device, rho, dx, dy, dz, 0., 0., 0., 0., dz, nz, False, False)
    lco, planeso = O.gtlchg3d(rho, dx, dy, dz, 0., 0., 0., 0., dz, nz)
    assert _bits_equal(planes, planeso) and _bits_equal(lc, lco)
    np.testing.assert_allclose(planes, rho.sum(axis=(1, 2)), rtol=1e-12, atol=0)
    e = _getese_gpu(device, rho, phi, dx, dy, dz, False, False)
    assert _bits_equal(e, O.getese3d(rho, phi, dx, dy, dz))
    want = 0.5 * dx * dy * dz * float(np.sum(rho * phi[1:-1, 1:-1, 1:-1]))
    assert e == pytest.approx(want, rel=1e-12)


def test_bad_input_is_an_error(device):
    from warp_amd import _native as N
    r = torch.zeros(8, dtype=torch.float64, device=device)
    with pytest.raises(RuntimeError, match="negative mesh size"):
        N.call("wm_gtlchg3d", r.data_ptr(), -1, 1, 1, 1., 1., 1., 0., 0., 0., 0., 1., 1, 0, 0,
               r.data_ptr(), r.data_ptr(), N.stream_of(device))
    with pytest.raises(RuntimeError, match="NULL"):
        N.call("wm_getese3d", r.data_ptr(), None, 1, 1, 1, 1., 1., 1., 0, 0, r.data_ptr(),
               r.data_ptr(), N.stream_of(device))


def test_solver_diagnostics(device):
    """MultiGrid3D.gtlchg / getese fill top.linechg and top.ese from the
    solver's rho and phi, the Z_arrays mesh defaulting to the field mesh
    (w3d.F:337-341)."""
    import warp_amd
    from warp_amd.globals import top
    saved = {k: getattr(top, k) for k in ("nzzarr", "zzmin", "zzmax", "dzz", "linechg", "ese")}
    try:
        top.nzzarr, top.zzmin, top.zzmax, top.dzz, top.linechg = 0, 0., 0., 0., None
        nx, ny, nz = 32, 24, 40
        s = warp_amd.MultiGrid3D(nx=nx, ny=ny, nz=nz, xmmin=-0.02, xmmax=0.02, ymmin=-0.015,
                                 ymmax=0.015, zmmin=0.01, zmmax=0.09, bounds=[0] * 6,
                                 mgverbose=-1)
        rng = np.random.default_rng(2)
        s.rho = _t(rng.uniform(0., 1e-7, (nz + 1, ny + 1, nx + 1)), device)
        s.solve()
        s.gtlchg()
        s.getese()
        rho, phi = s.rho.cpu().numpy(), s.phi.cpu().numpy()
        assert top.nzzarr == nz and top.zzmin == 0.01 and top.zzmax == 0.09
        lco, _ = O.gtlchg3d(rho, s.dx, s.dy, s.dz, top.zgrid, s.zmmin, top.zzmin, top.zbeam,
                            top.dzz, top.nzzarr)
        assert _bits_equal(top.linechg.cpu().numpy(), lco)
        assert _bits_equal(top.ese, O.getese3d(rho, phi, s.dx, s.dy, s.dz))
        assert top.ese > 0.
    finally:
        for k, v in saved.items():
            setattr(top, k, v)
