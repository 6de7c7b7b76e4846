/*
 * oc_particles.c -- TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline).
 *
 * FP64 restatement of the particle side of the reference's electrostatic
 * PIC step: CIC deposit (setrho3ddirect1), rho boundary folding, field
 * gather from phi (sete_from_phi_linear3d), grid E (getselfe3d), and the
 * leapfrog kicks/drift (epush3d, xpush3d), from pywarp90/w3d.F.  No
 * symmetries (l2symtry = l4symtry = .false.), ny > 0, nxguardphi = 1,
 * nxguardrho = nxguarde = 0.  Compile with -ffp-contract=off.
 */
#include "warp_oracle.h"

#define IDXP(ix, iy, iz) \
  ((int64_t)((ix) + 1) + (nx + 3) * ((int64_t)((iy) + 1) + (ny + 3) * (int64_t)((iz) + 1)))
#define IDXR(ix, iy, iz) \
  ((int64_t)(ix) + (nx + 1) * ((int64_t)(iy) + (ny + 1) * (int64_t)(iz)))

/* setrho3ddirect1, w3d.F:7757-7963, "normal loop" branch (7843-7916).
 * The one-particle-ahead prefetch of the reference does not change any
 * value, so each particle's index and weights are computed directly. */
void oc_setrho3d(double *rho, int64_t np, const double *xp, const double *yp,
                 const double *zp, double zgrid, double q, double wght,
                 int64_t nx, int64_t ny, int64_t nz,
                 double dx, double dy, double dz,
                 double xmmin, double ymmin, double zmmin) {
  double g = wght * q;
  if (nx > 0) g = g / dx;
  if (ny > 0) g = g / dy;
  if (nz > 0) g = g / dz;
  double dxi = 1. / dx, dyi = 1. / dy, dzi = 1. / dz;
  int64_t ip;
  for (ip = 0; ip < np; ip++) {
    int64_t ii = (int64_t)((xp[ip] - xmmin) * dxi);
    int64_t jj = (int64_t)((yp[ip] - ymmin) * dyi);
    int64_t kk = (int64_t)((zp[ip] - zgrid - zmmin) * dzi);
    double u1 = (xp[ip] - xmmin) * dxi - ii;
    double v1 = (yp[ip] - ymmin) * dyi - jj;
    double w1 = (zp[ip] - zgrid - zmmin) * dzi - kk;
    double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
    rho[IDXR(ii, jj, kk)] = rho[IDXR(ii, jj, kk)] + u0 * v0 * w0 * g;
    rho[IDXR(ii + 1, jj, kk)] = rho[IDXR(ii + 1, jj, kk)] + u1 * v0 * w0 * g;
    rho[IDXR(ii, jj + 1, kk)] = rho[IDXR(ii, jj + 1, kk)] + u0 * v1 * w0 * g;
    rho[IDXR(ii + 1, jj + 1, kk)] = rho[IDXR(ii + 1, jj + 1, kk)] + u1 * v1 * w0 * g;
    rho[IDXR(ii, jj, kk + 1)] = rho[IDXR(ii, jj, kk + 1)] + u0 * v0 * w1 * g;
    rho[IDXR(ii + 1, jj, kk + 1)] = rho[IDXR(ii + 1, jj, kk + 1)] + u1 * v0 * w1 * g;
    rho[IDXR(ii, jj + 1, kk + 1)] = rho[IDXR(ii, jj + 1, kk + 1)] + u0 * v1 * w1 * g;
    rho[IDXR(ii + 1, jj + 1, kk + 1)] = rho[IDXR(ii + 1, jj + 1, kk + 1)] + u1 * v1 * w1 * g;
  }
}

/* applyrhoboundaryconditions3d, w3d.F:2660-2750, serial, nxguardrho = 0:
 * Neumann faces double, periodic faces sum and copy. */
void oc_applyrhoboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz,
                                     double *rho, const int64_t *b) {
  int64_t ix, iy, iz;
  if (nx > 0) {
    for (iz = 0; iz <= nz; iz++)
      for (iy = 0; iy <= ny; iy++) {
        if (b[0] == 1) rho[IDXR(0, iy, iz)] = rho[IDXR(0, iy, iz)] + rho[IDXR(0, iy, iz)];
        if (b[1] == 1) rho[IDXR(nx, iy, iz)] = rho[IDXR(nx, iy, iz)] + rho[IDXR(nx, iy, iz)];
        if (b[0] == 2 || b[1] == 2) {
          rho[IDXR(0, iy, iz)] = rho[IDXR(0, iy, iz)] + rho[IDXR(nx, iy, iz)];
          rho[IDXR(nx, iy, iz)] = rho[IDXR(0, iy, iz)];
        }
      }
  }
  if (ny > 0) {
    for (iz = 0; iz <= nz; iz++)
      for (ix = 0; ix <= nx; ix++) {
        if (b[2] == 1) rho[IDXR(ix, 0, iz)] = rho[IDXR(ix, 0, iz)] + rho[IDXR(ix, 0, iz)];
        if (b[3] == 1) rho[IDXR(ix, ny, iz)] = rho[IDXR(ix, ny, iz)] + rho[IDXR(ix, ny, iz)];
        if (b[2] == 2 || b[3] == 2) {
          rho[IDXR(ix, 0, iz)] = rho[IDXR(ix, 0, iz)] + rho[IDXR(ix, ny, iz)];
          rho[IDXR(ix, ny, iz)] = rho[IDXR(ix, 0, iz)];
        }
      }
  }
  if (nz > 0) {
    for (iy = 0; iy <= ny; iy++)
      for (ix = 0; ix <= nx; ix++) {
        if (b[4] == 1) rho[IDXR(ix, iy, 0)] = rho[IDXR(ix, iy, 0)] + rho[IDXR(ix, iy, 0)];
        if (b[5] == 1) rho[IDXR(ix, iy, nz)] = rho[IDXR(ix, iy, nz)] + rho[IDXR(ix, iy, nz)];
        if (b[4] == 2 || b[5] == 2) {
          rho[IDXR(ix, iy, 0)] = rho[IDXR(ix, iy, 0)] + rho[IDXR(ix, iy, nz)];
          rho[IDXR(ix, iy, nz)] = rho[IDXR(ix, iy, 0)];
        }
      }
  }
}

/* sete_from_phi_linear3d, w3d.F:3963-4107 (no symmetry).  E is added to
 * the incoming ex, ey, ez. */
void oc_sete3d(const double *phi, int64_t np, const double *xp,
               const double *yp, const double *zp, double zgrid,
               double xmmin, double ymmin, double zmmin,
               double dx, double dy, double dz,
               int64_t nx, int64_t ny, int64_t nz,
               double *ex, double *ey, double *ez) {
  double tdxi = 1. / (2. * dx), tdyi = 1. / (2. * dy), tdzi = 1. / (2. * dz);
  double dxi = 1. / dx, dyi = 1. / dy, dzi = 1. / dz;
  int64_t ip;
#define P(a, b, c) phi[IDXP(i + (a), j + (b), k + (c))]
  OC_PARFOR()
  for (ip = 0; ip < np; ip++) {
    int64_t i = (int64_t)((xp[ip] - xmmin) * dxi);
    int64_t j = (int64_t)((yp[ip] - ymmin) * dyi);
    int64_t k = (int64_t)((zp[ip] - zgrid - zmmin) * dzi);
    double u1 = (xp[ip] - xmmin) * dxi - i;
    double v1 = (yp[ip] - ymmin) * dyi - j;
    double w1 = (zp[ip] - zgrid - zmmin) * dzi - k;
    double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
    double ext = tdxi * (u0 * v0 * w0 * (P(-1, 0, 0) - P(1, 0, 0)) +
                         u1 * v0 * w0 * (P(0, 0, 0) - P(2, 0, 0)) +
                         u0 * v1 * w0 * (P(-1, 1, 0) - P(1, 1, 0)) +
                         u1 * v1 * w0 * (P(0, 1, 0) - P(2, 1, 0)) +
                         u0 * v0 * w1 * (P(-1, 0, 1) - P(1, 0, 1)) +
                         u1 * v0 * w1 * (P(0, 0, 1) - P(2, 0, 1)) +
                         u0 * v1 * w1 * (P(-1, 1, 1) - P(1, 1, 1)) +
                         u1 * v1 * w1 * (P(0, 1, 1) - P(2, 1, 1)));
    double eyt = tdyi * (u0 * v0 * w0 * (P(0, -1, 0) - P(0, 1, 0)) +
                         u1 * v0 * w0 * (P(1, -1, 0) - P(1, 1, 0)) +
                         u0 * v1 * w0 * (P(0, 0, 0) - P(0, 2, 0)) +
                         u1 * v1 * w0 * (P(1, 0, 0) - P(1, 2, 0)) +
                         u0 * v0 * w1 * (P(0, -1, 1) - P(0, 1, 1)) +
                         u1 * v0 * w1 * (P(1, -1, 1) - P(1, 1, 1)) +
                         u0 * v1 * w1 * (P(0, 0, 1) - P(0, 2, 1)) +
                         u1 * v1 * w1 * (P(1, 0, 1) - P(1, 2, 1)));
    double ezt = tdzi * (u0 * v0 * w0 * (P(0, 0, -1) - P(0, 0, 1)) +
                         u1 * v0 * w0 * (P(1, 0, -1) - P(1, 0, 1)) +
                         u0 * v1 * w0 * (P(0, 1, -1) - P(0, 1, 1)) +
                         u1 * v1 * w0 * (P(1, 1, -1) - P(1, 1, 1)) +
                         u0 * v0 * w1 * (P(0, 0, 0) - P(0, 0, 2)) +
                         u1 * v0 * w1 * (P(1, 0, 0) - P(1, 0, 2)) +
                         u0 * v1 * w1 * (P(0, 1, 0) - P(0, 1, 2)) +
                         u1 * v1 * w1 * (P(1, 1, 0) - P(1, 1, 2)));
    ex[ip] = ex[ip] + ext;
    ey[ip] = ey[ip] + eyt;
    ez[ip] = ez[ip] + ezt;
  }
#undef P
}

/* epush3d, w3d.F:2752-2777. */
void oc_epush3d(int64_t np, double *uxp, double *uyp, double *uzp,
                const double *ex, const double *ey, const double *ez,
                double q, double m, double ddt) {
  double cst = q * ddt / m;
  int64_t ip;
  OC_PARFOR()
  for (ip = 0; ip < np; ip++) {
    uxp[ip] = uxp[ip] + ex[ip] * cst;
    uyp[ip] = uyp[ip] + ey[ip] * cst;
    uzp[ip] = uzp[ip] + ez[ip] * cst;
  }
}

/* xpush3d, w3d.F:3070-3092. */
void oc_xpush3d(int64_t np, double *xp, double *yp, double *zp,
                const double *uxp, const double *uyp, const double *uzp,
                const double *gaminv, double dtp) {
  int64_t ip;
  OC_PARFOR()
  for (ip = 0; ip < np; ip++) {
    xp[ip] = xp[ip] + uxp[ip] * gaminv[ip] * dtp;
    yp[ip] = yp[ip] + uyp[ip] * gaminv[ip] * dtp;
    zp[ip] = zp[ip] + uzp[ip] * gaminv[ip] * dtp;
  }
}

/* getselfe3d, w3d.F:6023-6109, nxguarde = 0: selfe(0:2,0:nx,0:ny,0:nz). */
void oc_getselfe3d(const double *phi, int64_t nx, int64_t ny, int64_t nz,
                   double *selfe, double dx, double dy, double dz,
                   int64_t lzero) {
  double tdxi = 0.5 / dx, tdyi = 0.5 / dy, tdzi = 0.5 / dz;
  int64_t ix, iy, iz;
  OC_PARFOR(private(iy, ix))
  for (iz = 0; iz <= nz; iz++)
    for (iy = 0; iy <= ny; iy++)
      for (ix = 0; ix <= nx; ix++) {
        double *e = selfe + 3 * IDXR(ix, iy, iz);
        double gx = tdxi * (phi[IDXP(ix - 1, iy, iz)] - phi[IDXP(ix + 1, iy, iz)]);
        double gy = tdyi * (phi[IDXP(ix, iy - 1, iz)] - phi[IDXP(ix, iy + 1, iz)]);
        double gz = tdzi * (phi[IDXP(ix, iy, iz - 1)] - phi[IDXP(ix, iy, iz + 1)]);
        if (lzero) {
          e[0] = gx; e[1] = gy; e[2] = gz;
        } else {
          e[0] = e[0] + gx; e[1] = e[1] + gy; e[2] = e[2] + gz;
        }
      }
}

/* sete_from_e_linear3d, w3d.F:4362-4460 (no symmetry), nxguarde = 0:
 * E at each particle by trilinear interpolation of the precomputed grid
 * field selfe(3, 0:nx, 0:ny, 0:nz), added to ex, ey, ez.  Particles with
 * a scaled position outside [0, n] are skipped (w3d.F:4425-4427).  For a
 * particle exactly on the upper face (i = n) the reference reads selfe at
 * n+1 with a zero weight; here that term reads selfe at n (same zero
 * contribution for finite fields). */
void oc_sete_from_e3d(const double *selfe, int64_t np, const double *xp,
                      const double *yp, const double *zp, double zgrid,
                      double xmmin, double ymmin, double zmmin, double dx,
                      double dy, double dz, int64_t nx, int64_t ny, int64_t nz,
                      double *ex, double *ey, double *ez) {
  double dxi = 1. / dx, dyi = 1. / dy, dzi = 1. / dz;
  int64_t ip;
  OC_PARFOR()
  for (ip = 0; ip < np; ip++) {
    double xi = (xp[ip] - xmmin) * dxi;
    double yj = (yp[ip] - ymmin) * dyi;
    double zk = (zp[ip] - zgrid - zmmin) * dzi;
    if (!(xi >= 0. && xi <= nx && yj >= 0. && yj <= ny && zk >= 0. && zk <= nz)) continue;
    int64_t i = (int64_t)xi, j = (int64_t)yj, k = (int64_t)zk;
    int64_t i1 = i < nx ? i + 1 : i, j1 = j < ny ? j + 1 : j, k1 = k < nz ? k + 1 : k;
    double u1 = xi - i, v1 = yj - j, w1 = zk - k;
    double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
    double out[3];
    int c;
#define E(a, b, d) selfe[3 * IDXR(a, b, d) + c]
    for (c = 0; c < 3; c++)
      out[c] = u0 * v0 * w0 * E(i, j, k) + u1 * v0 * w0 * E(i1, j, k) +
               u0 * v1 * w0 * E(i, j1, k) + u1 * v1 * w0 * E(i1, j1, k) +
               u0 * v0 * w1 * E(i, j, k1) + u1 * v0 * w1 * E(i1, j, k1) +
               u0 * v1 * w1 * E(i, j1, k1) + u1 * v1 * w1 * E(i1, j1, k1);
#undef E
    ex[ip] = ex[ip] + out[0];
    ey[ip] = ey[ip] + out[1];
    ez[ip] = ez[ip] + out[2];
  }
}

/* getzmmnt, top.F:521-1360, itask == 2, the serial (non-J90) sums of the
 * whole-beam moments zmmnts0(1:33) (top.F:1059-1136).  usegaminv selects
 * the branch for relativistic velocities without momentum moments
 * (lrelativ and not l_momentum_moments).  mm[k-1] holds zmmnts0(k). */
void oc_getzmmnt(int64_t np, const double *xp, const double *yp, const double *zp,
                 const double *uxp, const double *uyp, const double *uzp,
                 const double *gaminv, int64_t usegaminv, double *mm) {
  const double smallpos = 1.0e-36;
  int64_t ip;
  for (ip = 0; ip < np; ip++) {
    double x = xp[ip], y = yp[ip], z = zp[ip], ux = uxp[ip], uy = uyp[ip], uz = uzp[ip];
    double vzi = 1. / (uz + smallpos);
    mm[0] = mm[0] + 1.;
    mm[1] = mm[1] + x;
    mm[2] = mm[2] + y;
    mm[3] = mm[3] + z;
    if (!usegaminv) {
      mm[6] = mm[6] + ux;
      mm[7] = mm[7] + uy;
      mm[8] = mm[8] + uz;
      mm[9] = mm[9] + x * y;
      mm[11] = mm[11] + x * x;
      mm[12] = mm[12] + y * y;
      mm[13] = mm[13] + z * z;
      mm[16] = mm[16] + ux * ux;
      mm[17] = mm[17] + uy * uy;
      mm[18] = mm[18] + uz * uz;
      mm[21] = mm[21] + z * uz;
      mm[24] = mm[24] + x * uz;
      mm[25] = mm[25] + y * uz;
      mm[26] = mm[26] + ux * uz;
      mm[27] = mm[27] + uy * uz;
    } else {
      double g = gaminv[ip];
      mm[6] = mm[6] + ux * g;
      mm[7] = mm[7] + uy * g;
      mm[8] = mm[8] + uz * g;
      mm[9] = mm[9] + x * y;
      mm[11] = mm[11] + x * x;
      mm[12] = mm[12] + y * y;
      mm[13] = mm[13] + z * z;
      mm[16] = mm[16] + (g * ux) * (g * ux);
      mm[17] = mm[17] + (g * uy) * (g * uy);
      mm[18] = mm[18] + (g * uz) * (g * uz);
      mm[21] = mm[21] + z * g * uz;
      mm[24] = mm[24] + x * uz * g;
      mm[25] = mm[25] + y * uz * g;
      mm[26] = mm[26] + ux * uz * (g * g);
      mm[27] = mm[27] + uy * uz * (g * g);
    }
    mm[4] = mm[4] + ux * vzi;
    mm[5] = mm[5] + uy * vzi;
    mm[10] = mm[10] + ux * uy * (vzi * vzi);
    mm[14] = mm[14] + (ux * vzi) * (ux * vzi);
    mm[15] = mm[15] + (uy * vzi) * (uy * vzi);
    mm[19] = mm[19] + x * ux * vzi;
    mm[20] = mm[20] + y * uy * vzi;
    mm[22] = mm[22] + x * uy * vzi;
    mm[23] = mm[23] + y * ux * vzi;
    if (!usegaminv) {
      mm[28] = mm[28] + x * ux;
      mm[29] = mm[29] + y * uy;
      mm[30] = mm[30] + x * uy;
      mm[31] = mm[31] + y * ux;
      mm[32] = mm[32] + ux * uy;
    } else {
      double g = gaminv[ip];
      mm[28] = mm[28] + x * ux * g;
      mm[29] = mm[29] + y * uy * g;
      mm[30] = mm[30] + x * uy * g;
      mm[31] = mm[31] + y * ux * g;
      mm[32] = mm[32] + ux * uy * (g * g);
    }
  }
}

/* The absorbing particle boundaries of particleboundaries3d
 * (w3d.F:11292-11450) with pboundxy = pbound0 = pboundnz = absorb and
 * lcountaslost: xparticleboundaries (top.F:3499-3580) and its y twin flag a
 * particle with gaminv = 0 when xp <= xmin or xp >= xmax; zparticleboundaries
 * (top.F:3735-3800) when zp-zg >= zmax or zp-zg <= zmin.  The extents are
 * those getparticleextent and particleboundaries3d pass: xmax = xmmin+nx*dx,
 * zmax = zmmin+nz*dz (w3d.F:11344-11356). */
void oc_particleboundaries3d(int64_t np, const double *xp, const double *yp,
                             double *zp, double *gaminv, double xmmin,
                             double ymmin, double zmmin, double dx, double dy,
                             double dz, int64_t nx, int64_t ny, int64_t nz,
                             double zgrid, int64_t zperiodic) {
  double xmax = xmmin + nx * dx, ymax = ymmin + ny * dy, zmax = zmmin + nz * dz;
  int64_t ip;
  for (ip = 0; ip < np; ip++)
    if (xp[ip] <= xmmin || xp[ip] >= xmax) gaminv[ip] = 0.;
  for (ip = 0; ip < np; ip++)
    if (yp[ip] <= ymmin || yp[ip] >= ymax) gaminv[ip] = 0.;
  if (zperiodic) {
    /* pbound0 = periodic, top.F:3762-3778 */
    double syslen = zmax - zmmin;
    for (ip = 0; ip < np; ip++) {
      if (zp[ip] < zgrid + zmmin)
        zp[ip] = zp[ip] + syslen * (int64_t)((zgrid + zmax - zp[ip]) / syslen);
      if (zp[ip] >= zgrid + zmax)
        zp[ip] = zp[ip] - syslen * (int64_t)((zp[ip] - zgrid - zmmin) / syslen);
    }
    return;
  }
  for (ip = 0; ip < np; ip++) {
    if ((zp[ip] - zgrid) >= zmax)
      gaminv[ip] = 0.;
    else if ((zp[ip] - zgrid) <= zmmin)
      gaminv[ip] = 0.;
  }
}

/* clearpart (top.F:4730-4893) on one species block of np particles whose
 * columns are cols[0..ncol); cols[iuz] is uzp and cols[igam] gaminv (the
 * lost flag: gaminv <= 0).  fillmethod 1 moves live particles from the end
 * into the holes, 2 shifts the live particles down keeping their order.
 * Returns the new nps. */
int64_t oc_clearpart(int64_t np, int64_t ncol, double **cols, int64_t iuz,
                     int64_t igam, int64_t fillmethod) {
  double *g = cols[igam];
  int64_t nps = 0, c;
  if (fillmethod == 1) {
    int64_t ips = np - 1, ip = 0;
    while (ip <= ips) {
      if (g[ip] <= 0.) {
        while (g[ips] <= 0. && ips > ip) ips--;
        if (ips > ip) {
          for (c = 0; c < ncol; c++) cols[c][ip] = cols[c][ips];
          cols[iuz][ips] = 0.;
          g[ips] = 0.;
          ips--;
          nps++;
        }
      } else {
        nps++;
      }
      ip++;
    }
  } else if (fillmethod == 2) {
    int64_t ip = 0, ips;
    for (ips = 0; ips < np; ips++) {
      if (g[ips] > 0.) {
        if (ips > ip) {
          for (c = 0; c < ncol; c++) cols[c][ip] = cols[c][ips];
          cols[iuz][ips] = 0.;
          g[ips] = 0.;
        }
        ip++;
        nps++;
      }
    }
  } else {
    nps = np;
  }
  return nps;
}

/* getgrid3d (dtop.F:2818-2920), ny > 0 branch: linear interpolation of the
 * node-centred grid (0:nx,0:ny,0:nz) over [xmin,xmax]x[ymin,ymax]x[zmin,zmax]
 * at the positions; f is left unchanged outside the grid.  The fetchphi
 * path passes phi with its guard cells (multigrid.py:560-575). */
void oc_getgrid3d(int64_t np, const double *x, const double *y, const double *z,
                  double *f, int64_t nx, int64_t ny, int64_t nz, const double *grid,
                  double xmin, double xmax, double ymin, double ymax, double zmin,
                  double zmax, int64_t l2symtry, int64_t l4symtry) {
  double dxi = nx / (xmax - xmin);
  double dyi = (ymax == ymin) ? 1. : ny / (ymax - ymin);
  double dzi = nz / (zmax - zmin);
  int64_t ip;
#define G(i, j, k) grid[(i) + (nx + 1) * ((j) + (ny + 1) * (int64_t)(k))]
  for (ip = 0; ip < np; ip++) {
    double xx, yy, gx = 0., gy = 0., gz = 0.;
    if (l4symtry) {
      xx = fabs(x[ip] - 0.) + 0.;
      yy = fabs(y[ip] - 0.) + 0.;
    } else if (l2symtry) {
      xx = x[ip];
      yy = fabs(y[ip] - 0.) + 0.;
    } else {
      xx = x[ip];
      yy = y[ip];
    }
    if (nx > 0) gx = (xx - xmin) * dxi;
    if (ny > 0) gy = (yy - ymin) * dyi;
    if (nz > 0) gz = (z[ip] - zmin) * dzi;
    if (0. <= gx && gx < nx && 0. <= gy && gy < ny && 0. <= gz && gz < nz) {
      int64_t ix = (int64_t)gx, iy = (int64_t)gy, iz = (int64_t)gz;
      double wx = gx - ix, wy = gy - iy, wz = gz - iz;
      f[ip] = G(ix, iy, iz) * (1. - wx) * (1. - wy) * (1. - wz) +
              G(ix + 1, iy, iz) * wx * (1. - wy) * (1. - wz) +
              G(ix, iy + 1, iz) * (1. - wx) * wy * (1. - wz) +
              G(ix + 1, iy + 1, iz) * wx * wy * (1. - wz) +
              G(ix, iy, iz + 1) * (1. - wx) * (1. - wy) * wz +
              G(ix + 1, iy, iz + 1) * wx * (1. - wy) * wz +
              G(ix, iy + 1, iz + 1) * (1. - wx) * wy * wz +
              G(ix + 1, iy + 1, iz + 1) * wx * wy * wz;
    }
  }
#undef G
}
