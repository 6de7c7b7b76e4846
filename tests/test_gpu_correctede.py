"""The conductor-corrected E field (f3d.lcorrectede) on the GPU against
the FP64 oracle, bit for bit: the node classification, the one-sided
subgrid E and its extrapolation into the conductor, the doubled grid E at
interior nodes, the gather for efetch 1 and 3 and the fused advance; on
slabs with a closed-form answer and on the pipe + annulus geometry."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

DOM = (-0.02, 0.02, -0.02, 0.02, 0.0, 0.08)
N3 = (32, 32, 64)
D3 = (0.04 / 32, 0.04 / 32, 0.08 / 64)
LO = (DOM[0], DOM[2], DOM[4])


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


class DevCondField:
    """A wm_condfield made from a lists() dict (None: no conductors)."""

    def __init__(self, lists, n3):
        from warp_amd import _native as N
        self.N = N
        self.n3 = n3
        self.view = H.lists_struct(lists) if lists is not None else None
        h = C.c_void_p()
        N.call("wm_condfield_create", C.byref(h), *n3,
               C.addressof(self.view) if self.view is not None else None)
        self.h = h

    def close(self):
        self.N.lib().wm_condfield_destroy(self.h)

    def read(self, device):
        N = self.N
        nsub = (C.c_int64 * 2)()
        N.call("wm_condfield_read", self.h, nsub, None, None, None, None, None, None)
        nx, ny, nz = self.n3
        icg = torch.empty((nz + 1, ny + 1, nx + 1), dtype=torch.int32, device=device)
        arr = [torch.empty((nsub[p], w), dtype=torch.float64, device=device)
               for p in range(2) for w in (3, 6)]
        N.call("wm_condfield_read", self.h, nsub, icg.data_ptr(),
               *[a.data_ptr() if a.numel() else None for a in arr], N.stream_of(device))
        a = [x.cpu().numpy() for x in arr]
        return icg.cpu().numpy(), [dict(efield0=a[0], efieldd=a[1]), dict(efield0=a[2], efieldd=a[3])]


def _check_arrays(dcf, ocf, device):
    icg, sub = dcf.read(device)
    np.testing.assert_array_equal(icg, ocf.icgrid)
    for p in range(2):
        for k in ("efield0", "efieldd"):
            np.testing.assert_array_equal(sub[p][k], ocf.sub[p][k])


def _fields(dcf, phi, efetch, n3, d3, device):
    """(device phi, device selfe or None, oracle selfe or None) after the
    efield and fixefield steps on both sides."""
    from warp_amd import _native as N
    st = N.stream_of(device)
    pd = _t(phi, device)
    N.call("wm_condfield_efield", dcf.h, pd.data_ptr(), *d3, st)
    if efetch != 3:
        return pd, None
    nx, ny, nz = n3
    sd = torch.empty((nz + 1, ny + 1, nx + 1, 3), dtype=torch.float64, device=device)
    N.call("wm_getselfe3d", pd.data_ptr(), nx, ny, nz, sd.data_ptr(), *d3, 1, st)
    N.call("wm_condfield_fixefield", dcf.h, sd.data_ptr(), st)
    return pd, sd


def _gather(dcf, pd, sd, efetch, pos, e0, lo, d3, device):
    from warp_amd import _native as N
    xd = [_t(p, device) for p in pos]
    ed = [_t(e, device) for e in e0]
    N.call("wm_sete3d_conductor", dcf.h, pd.data_ptr(), sd.data_ptr() if sd is not None else None,
           efetch, len(pos[0]), *[a.data_ptr() for a in xd], 0.0, *lo, *d3,
           *[a.data_ptr() for a in ed], N.stream_of(device))
    return [a.cpu().numpy() for a in ed]


@pytest.mark.parametrize("efetch", [1, 3])
@pytest.mark.parametrize("axis,side,frac", [(0, -1, 0.4), (0, 1, 0.85), (1, -1, 1.0), (1, 1, 0.4),
                                            (2, -1, 0.85), (2, 1, 1.0)])
def test_slab_bitwise_and_exact(device, axis, side, frac, efetch):
    """Slabs normal to each axis: the device arrays and gathered E equal
    the oracle's bit for bit, and the E is the closed-form field."""
    from tests.test_oracle import SLAB_COEF, SLAB_D3, SLAB_N3, slab_case, slab_expected
    ocf, phi, selfe, pos, eo = slab_case(axis, side, frac, efetch)
    d = H.slab_lists(SLAB_N3, axis, side, 4, frac, 3.0)
    dcf = DevCondField(d, SLAB_N3)
    try:
        pd, sd = _fields(dcf, phi, efetch, SLAB_N3, SLAB_D3, device)
        _check_arrays(dcf, ocf, device)
        if efetch == 3:
            np.testing.assert_array_equal(sd.cpu().numpy(), selfe)
        e = _gather(dcf, pd, sd, efetch, pos, [np.zeros(len(pos[0]))] * 3, (0., 0., 0.),
                    SLAB_D3, device)
    finally:
        dcf.close()
    want, _, _ = slab_expected(axis, side, frac, pos)
    for c in range(3):
        np.testing.assert_array_equal(e[c], eo[c])
        np.testing.assert_allclose(e[c], want[c], rtol=0,
                                   atol=1e-12 * (np.abs(want[c]).max() + 1e-300))


def _pipe_annulus_lists():
    co = O.install_conductors(H.pipe_and_annulus(O), *N3, *DOM)
    O.checkconductors(co, *N3, *D3)
    return co.lists()


def _particles(n, seed):
    """Uniform over the mesh (every cell kind), plus nodes, upper faces,
    points outside and a NaN."""
    rng = np.random.default_rng(seed)
    pos = [LO[a] + rng.uniform(0, N3[a] * D3[a], n) for a in range(3)]
    extra = np.array([[LO[0] + N3[0] * D3[0], 0.0, 0.04], [0.0, LO[1] + N3[1] * D3[1], 0.04],
                      [0.001, -0.012, N3[2] * D3[2]], [LO[0], LO[1], 0.0], [-0.03, 0.0, 0.04],
                      [0.0, 0.0, -1e-3], [np.nan, 0.0, 0.04], [0.0125, 0.0, 0.0325]])
    return [np.concatenate([pos[a], extra[:, a]]) for a in range(3)]


@pytest.mark.parametrize("efetch", [1, 3])
def test_pipe_annulus_bitwise(device, efetch):
    """The pipe + annulus lists (both parities, interior and subgrid nodes
    on curved surfaces): arrays and gather equal the oracle bit for bit;
    E accumulates onto what ex, ey, ez hold."""
    lists = _pipe_annulus_lists()
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((N3[2] + 3, N3[1] + 3, N3[0] + 3)) * 50.
    ocf = O.CondField(lists, *N3)
    ocf.efield(phi, *D3)
    se = None
    if efetch == 3:
        se = O.getselfe3d(phi, *D3)
        ocf.fixefield(se)
    pos = _particles(100000, 6)
    e0 = [np.full(len(pos[0]), 0.5) for _ in range(3)]
    eo = [a.copy() for a in e0]
    O.sete3d_conductor(ocf, phi, se, *pos, *eo, *D3, *LO, efetch=efetch)
    dcf = DevCondField(lists, N3)
    try:
        pd, sd = _fields(dcf, phi, efetch, N3, D3, device)
        _check_arrays(dcf, ocf, device)
        e = _gather(dcf, pd, sd, efetch, pos, e0, LO, D3, device)
    finally:
        dcf.close()
    near = ocf.icgrid != 0
    assert near.sum() > 1000
    for c in range(3):
        np.testing.assert_array_equal(e[c], eo[c])


@pytest.mark.parametrize("efetch,center", [(1, 0), (3, 0), (1, 1), (3, 2)])
def test_padvnc3d_conductor_bitwise(device, efetch, center):
    """The fused advance with the corrected E equals the oracle's
    E = 0, sete3dwithconductor, epush3d(dt/2) (twice for fullv),
    xpush3d(dt) (not for synchv) bit for bit, and counts the particles
    leaving the mesh."""
    from warp_amd import _native as N
    lists = _pipe_annulus_lists()
    rng = np.random.default_rng(7)
    phi = rng.standard_normal((N3[2] + 3, N3[1] + 3, N3[0] + 3)) * 1e3
    pos = _particles(80000, 8)
    pos = [np.nan_to_num(p, nan=0.0) for p in pos]
    n = len(pos[0])
    u = [rng.normal(0, 1e4, n), rng.normal(0, 1e4, n), 1e6 + rng.normal(0, 1e5, n)]
    q, m, dt = 1.602176565e-19, 39 * 1.660538921e-27, 1e-9
    ocf = O.CondField(lists, *N3)
    ocf.efield(phi, *D3)
    se = None
    if efetch == 3:
        se = O.getselfe3d(phi, *D3)
        ocf.fixefield(se)
    xo, uo = [p.copy() for p in pos], [a.copy() for a in u]
    eo = [np.zeros(n) for _ in range(3)]
    O.sete3d_conductor(ocf, phi, se, *xo, *eo, *D3, *LO, efetch=efetch)
    for _ in range(2 if center == 0 else 1):
        O.epush3d(*uo, *eo, q, m, 0.5 * (dt * 1 * 1.))
    if center != 2:
        O.xpush3d(*xo, *uo, np.ones(n), dt)
    out = ((xo[0] < DOM[0]) | (xo[0] >= DOM[1]) | (xo[1] < DOM[2]) | (xo[1] >= DOM[3]) |
           (xo[2] < DOM[4]) | (xo[2] >= DOM[5]))
    dcf = DevCondField(lists, N3)
    try:
        pd, sd = _fields(dcf, phi, efetch, N3, D3, device)
        g = [_t(a, device) for a in pos] + [_t(a, device) for a in u]
        gi = torch.ones(n, dtype=torch.float64, device=device)
        E = [torch.zeros(n, dtype=torch.float64, device=device) for _ in range(3)]
        nlost = torch.zeros(1, dtype=torch.int64, device=device)
        N.call("wm_padvnc3d_conductor", dcf.h, pd.data_ptr(),
               sd.data_ptr() if sd is not None else None, efetch, n,
               *[a.data_ptr() for a in g], gi.data_ptr(), *[a.data_ptr() for a in E], 0.0, *LO,
               *D3, q, m, dt, DOM[1], DOM[3], DOM[5], nlost.data_ptr(), center,
               N.stream_of(device))
        got = [a.cpu().numpy() for a in g + E]
        nl = int(nlost.item())
    finally:
        dcf.close()
    for a, b in zip(got, xo + uo + eo):
        np.testing.assert_array_equal(a, b)
    assert nl == (0 if center == 2 else int(out.sum()))


def test_no_conductors_is_vacuum_path(device):
    """A handle made without lists classifies every node as vacuum: the
    gather is the oracle's vacuum path (per-term 1/(2 dx))."""
    rng = np.random.default_rng(9)
    phi = rng.standard_normal((N3[2] + 3, N3[1] + 3, N3[0] + 3))
    empty = dict(levels=1, int_indx=np.zeros((0, 3), np.int64), int_volt=np.zeros(0),
                 int_istart=np.zeros(2, np.int64))
    for p in range(2):
        empty["sub%d_indx" % p] = np.zeros((0, 3), np.int64)
        empty["sub%d_dels" % p] = np.zeros((0, 6))
        empty["sub%d_volt" % p] = np.zeros((0, 6))
        empty["sub%d_istart" % p] = np.zeros(2, np.int64)
    ocf = O.CondField(empty, *N3)
    pos = _particles(20000, 10)
    eo = [np.zeros(len(pos[0])) for _ in range(3)]
    O.sete3d_conductor(ocf, phi, None, *pos, *eo, *D3, *LO, efetch=1)
    dcf = DevCondField(None, N3)
    try:
        pd, _ = _fields(dcf, phi, 1, N3, D3, device)
        _check_arrays(dcf, ocf, device)
        e = _gather(dcf, pd, None, 1, pos, [np.zeros(len(pos[0]))] * 3, LO, D3, device)
    finally:
        dcf.close()
    for c in range(3):
        np.testing.assert_array_equal(e[c], eo[c])


def test_condfield_rejects_bad_input(device):
    from warp_amd import _native as N
    d = H.slab_lists((4, 4, 4), 0, -1, 2, 0.5, 1.0)
    d["sub0_indx"][0, 0] = 9                       # outside the 4x4x4 grid
    v = H.lists_struct(d)
    h = C.c_void_p()
    with pytest.raises(RuntimeError, match="outside the grid"):
        N.call("wm_condfield_create", C.byref(h), 4, 4, 4, C.addressof(v))
    dcf = DevCondField(None, (4, 4, 4))
    try:
        with pytest.raises(RuntimeError, match="efetch"):
            N.call("wm_sete3d_conductor", dcf.h, None, None, 2, 0, None, None, None, 0.0,
                   0., 0., 0., 1., 1., 1., None, None, None, N.stream_of(device))
        with pytest.raises(RuntimeError, match="efetch 3 needs selfe"):
            N.call("wm_sete3d_conductor", dcf.h, None, None, 3, 0, None, None, None, 0.0,
                   0., 0., 0., 1., 1., 1., None, None, None, N.stream_of(device))
    finally:
        dcf.close()


@pytest.mark.parametrize("efetch", [1, 3])
def test_solver_lcorrectede_fetch(device, efetch):
    """MultiGrid3D under f3d.lcorrectede: after a solve with the pipe and
    annulus installed, fetchfieldfrompositions gives the oracle's
    sete3dwithconductor of the solver's phi and conductor lists bit for
    bit; with the flag off it is the plain gather (which differs near the
    conductors)."""
    import warp_amd
    from warp_amd.globals import f3d, top
    s = warp_amd.MultiGrid3D(nx=N3[0], ny=N3[1], nz=N3[2], xmmin=DOM[0], xmmax=DOM[1],
                             ymmin=DOM[2], ymmax=DOM[3], zmmin=DOM[4], zmmax=DOM[5],
                             bounds=[0] * 6, mgverbose=-1)
    for c in H.pipe_and_annulus(warp_amd):
        s.installconductor(c)
    z = DOM[4] + D3[2] * torch.arange(N3[2] + 1, dtype=torch.float64)
    y = DOM[2] + D3[1] * torch.arange(N3[1] + 1, dtype=torch.float64)
    x = DOM[0] + D3[0] * torch.arange(N3[0] + 1, dtype=torch.float64)
    Z, Y, X = torch.meshgrid(z, y, x, indexing="ij")
    saved = (f3d.lcorrectede, top.efetch)
    try:
        f3d.lcorrectede, top.efetch = True, efetch
        s.rho = (1e-7 * torch.exp(-(X ** 2 + Y ** 2) / 2e-5 - (Z - 0.04) ** 2 / 2e-4)).to(device)
        s.solve()
        phi = s.phi.cpu().numpy()
        ocf = O.CondField(s.getconductorobject().lists(), *N3)
        ocf.efield(phi, *D3)
        se = None
        if efetch == 3:
            se = O.getselfe3d(phi, *D3)
            ocf.fixefield(se)
        pos = _particles(50000, 11)
        eo = [np.zeros(len(pos[0])) for _ in range(3)]
        O.sete3d_conductor(ocf, phi, se, *pos, *eo, *D3, *LO, efetch=efetch)
        xd = [_t(p, device) for p in pos]
        ed = [torch.zeros(len(pos[0]), dtype=torch.float64, device=device) for _ in range(3)]
        s.fetchfieldfrompositions(*xd, *ed)
        for c in range(3):
            np.testing.assert_array_equal(ed[c].cpu().numpy(), eo[c])
        f3d.lcorrectede = False
        inmesh = [p[:50000] for p in xd]
        plain = [torch.zeros(50000, dtype=torch.float64, device=device) for _ in range(3)]
        s.fetchfieldfrompositions(*inmesh, *plain)
        diff = max(float((plain[c].cpu() - ed[c][:50000].cpu()).abs().max()) for c in range(3))
        assert diff > 1e-3 * float(np.abs(eo[0]).max())
    finally:
        f3d.lcorrectede, top.efetch = saved
