"""wm_getese3d_noalloc, the energy diagnostic whose column sums live in the
library's own device scratch instead of a caller's work buffer: bitwise
equal to wm_getese3d on consecutive calls (the second smaller than the
first, so stale sums of the first would show), in every symmetry mode, and
a clear error for meshes beyond the scratch."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bind():
    from warp_amd import _native as N
    L = N.lib()
    f = L.wm_getese3d_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                  C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
    return N, f


def test_noalloc_energy_equals_work_buffer_energy(device):
    N, f = _bind()
    rng = np.random.default_rng(12)
    s = N.stream_of(device)
    cases = [((300, 300, 40), 0, 0), ((64, 48, 80), 0, 0), ((37, 51, 9), 1, 0),
             ((20, 10, 33), 0, 1), ((5, 3, 2), 1, 1)]
    for (nx, ny, nz), l4, l2 in cases:
        rho = torch.from_numpy(rng.standard_normal((nz + 1, ny + 1, nx + 1))).to(device)
        phi = torch.from_numpy(rng.standard_normal((nz + 3, ny + 3, nx + 3))).to(device)
        work = torch.empty(int(N.lib().wm_getese3d_worksize(nx, ny)) // 8, dtype=torch.float64,
                           device=device)
        want = torch.zeros(1, dtype=torch.float64, device=device)
        got = torch.full((1,), np.nan, dtype=torch.float64, device=device)
        N.call("wm_getese3d", rho.data_ptr(), phi.data_ptr(), nx, ny, nz, 0.1, 0.2, 0.3, l4, l2,
               work.data_ptr(), want.data_ptr(), s)
        N.check(f(rho.data_ptr(), phi.data_ptr(), nx, ny, nz, 0.1, 0.2, 0.3, l4, l2,
                  got.data_ptr(), s), "wm_getese3d_noalloc")
        torch.cuda.synchronize()
        assert got.item() == want.item(), ((nx, ny, nz), l4, l2, got.item(), want.item())


def test_noalloc_energy_rejects_meshes_beyond_its_scratch(device):
    N, f = _bind()
    t = torch.zeros(8, dtype=torch.float64, device=device)
    rc = f(t.data_ptr(), t.data_ptr(), 400, 400, 4, 1., 1., 1., 0, 0, t.data_ptr(),
           N.stream_of(device))
    assert rc != 0
    msg = N.lib().wm_last_error().decode()
    assert "exceed" in msg and "131072" in msg
