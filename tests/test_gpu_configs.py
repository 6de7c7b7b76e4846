"""Each benchmark configuration (mesh, conductors and beam at full size)
through one PIC step on the GPU and with the FP64 oracle from the same
warm start: E gather and fullv advance, deposit at the new positions, field
solve to mgtol.  States the error of every precision mode against the FP64
oracle: particle positions after the step (bar 1e-10 relative), phi
(relative L2, bar 1e-6) and the grid E of phi (relative L2).  The FP64 path
differs from the oracle only through the order of the deposit's sums."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("name", ["c0", "c1", "c2", "c4"])
def test_config_step_against_oracle(device, name):
    from warp_amd import configs
    cfg = configs.CONFIGS[name]
    nx, ny, nz = cfg["n"]
    x0, x1, y0, y1, z0, z1 = configs.DOMAIN
    dx, dy, dz = (x1 - x0) / nx, (y1 - y0) / ny, (z1 - z0) / nz
    pg = configs.load_beam(cfg, device)
    s = configs.make_solver(cfg, "fp64", device, mgverbose=-1)
    # warm start: the solve of the loaded beam
    s.loadrho(lzero=True, pgroups=[pg])
    s.solve()
    phi0 = s.phi.cpu().numpy().copy()
    cols = [getattr(pg, c).cpu().numpy().copy() for c in ("xp", "yp", "zp", "uxp", "uyp", "uzp")]
    q, m, w = pg.sq[0], pg.sm[0], pg.sw[0]
    # one step on the GPU
    s.gather_push(pg, 0, configs.DT, "fullv")
    s.loadrho(lzero=True, pgroups=[pg])
    rho_in = s.rho.clone()
    s.solve()
    phi_gpu = s.phi.cpu().numpy()
    xg = [getattr(pg, c).cpu().numpy() for c in ("xp", "yp", "zp", "uxp", "uyp", "uzp")]
    # the same step with the oracle
    x, y, z, ux, uy, uz = [a.copy() for a in cols]
    n = len(x)
    e = [np.zeros(n) for _ in range(3)]
    O.sete3d(phi0, x, y, z, *e, dx, dy, dz, x0, y0, z0)
    O.epush3d(ux, uy, uz, *e, q, m, 0.5 * configs.DT)
    O.epush3d(ux, uy, uz, *e, q, m, 0.5 * configs.DT)
    O.xpush3d(x, y, z, ux, uy, uz, np.ones(n), configs.DT)
    pos_err = max(float(np.abs(a - b).max() / np.abs(b).max())
                  for a, b in zip(xg, [x, y, z, ux, uy, uz]))
    k = (x >= x0) & (x < x1) & (y >= y0) & (y < y1) & (z >= z0) & (z < z1)
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    O.setrho3d(rho, x[k], y[k], z[k], q, w, dx, dy, dz, x0, y0, z0)
    bounds = configs.field_bounds(cfg)
    O.applyrhoboundaryconditions3d(rho, bounds)
    co = O.install_conductors(configs.conductors(cfg["conductors"], O), nx, ny, nz, x0, x1,
                              y0, y1, z0, z1)
    phi_ref = phi0.copy()
    it, err = O.multigrid3dsolve(phi_ref, rho, dx, dy, dz, bounds=bounds, conductors=co)
    e_ref = O.getselfe3d(phi_ref, dx, dy, dz)
    report = {"fp64": (_rel_l2(phi_gpu, phi_ref),
                       _rel_l2(O.getselfe3d(phi_gpu, dx, dy, dz), e_ref), s.mgiters)}
    del s
    torch.cuda.empty_cache()
    for prec in ("fp32", "bf16"):
        sp = configs.make_solver(cfg, prec, device, mgverbose=-1)
        sp.phi = torch.from_numpy(phi0).to(device)
        sp.rho = rho_in
        sp.solve()
        p = sp.phi.cpu().numpy()
        report[prec] = (_rel_l2(p, phi_ref), _rel_l2(O.getselfe3d(p, dx, dy, dz), e_ref),
                        sp.mgiters)
        del sp
        torch.cuda.empty_cache()
    print("\n%s %dx%dx%d, %d particles: positions max rel err %.2e; oracle %d V-cycles; "
          % (name, nx, ny, nz, n, pos_err, it) +
          "; ".join("%s phi relL2 %.2e E relL2 %.2e (%d V-cycles)" % (kk, *v)
                    for kk, v in report.items()))
    assert pos_err <= 1e-10
    for kk, (ephi, ee, iters) in report.items():
        assert ephi <= 1e-6, (kk, ephi)
    assert report["fp64"][0] <= 1e-10
