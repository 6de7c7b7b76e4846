/*
 * oc_conductors.c -- TEST INFRASTRUCTURE ONLY (parity oracle).
 *
 * FP64 restatement of the reference's conductor set-up: grid-line
 * intercepts of z-aligned cylinders (ZCylinderConductorFnew,
 * ZCylinderOutConductorFnew) and of straight-segment surfaces of revolution
 * (ZSrfrvConductorFnew, used by ZAnnulus), and the conversion of intercepts
 * into interior points and subgrid distances (conductordelfromintercepts),
 * all from pywarp90/f3d_conductors.F.
 */
#include "warp_oracle.h"
#include <math.h>
#include <stdlib.h>

#define LARGEPOS 1.0e+36

#define XI(ii, iy, iz) icpt->xintercepts[((ii) - 1) + icpt->nxicpt * ((iy) + (ny + 1) * (iz))]
#define YI(ii, ix, iz) icpt->yintercepts[((ii) - 1) + icpt->nyicpt * ((ix) + (nx + 1) * (iz))]
#define ZI(ii, ix, iy) icpt->zintercepts[((ii) - 1) + icpt->nzicpt * ((ix) + (nx + 1) * (iy))]
#define XV(ii, iy, iz) icpt->xvoltages[((ii) - 1) + icpt->nxicpt * ((iy) + (ny + 1) * (iz))]
#define YV(ii, ix, iz) icpt->yvoltages[((ii) - 1) + icpt->nyicpt * ((ix) + (nx + 1) * (iz))]
#define ZV(ii, ix, iy) icpt->zvoltages[((ii) - 1) + icpt->nzicpt * ((ix) + (nx + 1) * (iy))]
#define XC(ii, iy, iz) icpt->xcondids[((ii) - 1) + icpt->nxicpt * ((iy) + (ny + 1) * (iz))]
#define YC(ii, ix, iz) icpt->ycondids[((ii) - 1) + icpt->nyicpt * ((ix) + (nx + 1) * (iz))]
#define ZC(ii, ix, iy) icpt->zcondids[((ii) - 1) + icpt->nzicpt * ((ix) + (nx + 1) * (iy))]

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* ZCylinderConductorFnew (f3d_conductors.F:2496-2583) when outside == 0,
 * ZCylinderOutConductorFnew (3158-3256) when outside != 0.  The caller has
 * set nxicpt = nyicpt = 2 (inside) or 4 (outside), nzicpt = 2. */
void oc_zcylinder_intercepts(double rad, double length, double xcent,
                             double ycent, double zcent, oc_intercepts *icpt,
                             double fuzz, int64_t outside) {
  double xmmin = icpt->xmmin, ymmin = icpt->ymmin, zmmin = icpt->zmmin;
  double dx = icpt->dx, dy = icpt->dy, dz = icpt->dz;
  int64_t nx = icpt->nx, ny = icpt->ny, nz = icpt->nz, ix, iy, iz, k;
  double zmin = -length * 0.5, zmax = +length * 0.5;
  for (iz = 0; iz <= nz; iz++) {
    double zz = zmmin + iz * dz - zcent;
    for (iy = 0; iy <= ny; iy++) {
      double yy = ymmin + iy * dy - ycent;
      if (!outside) {
        if (zmin - fuzz <= zz && zz <= zmax + fuzz && fabs(yy) <= rad) {
          XI(1, iy, iz) = xcent - sqrt(dmax(0., rad * rad - yy * yy));
          XI(2, iy, iz) = xcent + sqrt(dmax(0., rad * rad - yy * yy));
        } else {
          XI(1, iy, iz) = LARGEPOS;
          XI(2, iy, iz) = LARGEPOS;
        }
      } else {
        if (zmin - fuzz <= zz && zz <= zmax + fuzz) {
          if (fabs(yy) <= rad) {
            XI(1, iy, iz) = -LARGEPOS;
            XI(2, iy, iz) = xcent - sqrt(dmax(0., rad * rad - yy * yy));
            XI(3, iy, iz) = xcent + sqrt(dmax(0., rad * rad - yy * yy));
            XI(4, iy, iz) = +LARGEPOS;
          } else {
            XI(1, iy, iz) = -LARGEPOS;
            for (k = 2; k <= 4; k++) XI(k, iy, iz) = +LARGEPOS;
          }
        } else {
          for (k = 1; k <= 4; k++) XI(k, iy, iz) = +LARGEPOS;
        }
      }
    }
  }
  for (iz = 0; iz <= nz; iz++) {
    double zz = zmmin + iz * dz - zcent;
    for (ix = 0; ix <= nx; ix++) {
      double xx = xmmin + ix * dx - xcent;
      if (!outside) {
        if (zmin - fuzz <= zz && zz <= zmax + fuzz && fabs(xx) <= rad) {
          YI(1, ix, iz) = ycent - sqrt(dmax(0., rad * rad - xx * xx));
          YI(2, ix, iz) = ycent + sqrt(dmax(0., rad * rad - xx * xx));
        } else {
          YI(1, ix, iz) = LARGEPOS;
          YI(2, ix, iz) = LARGEPOS;
        }
      } else {
        if (zmin - fuzz <= zz && zz <= zmax + fuzz) {
          if (fabs(xx) <= rad) {
            YI(1, ix, iz) = -LARGEPOS;
            YI(2, ix, iz) = ycent - sqrt(dmax(0., rad * rad - xx * xx));
            YI(3, ix, iz) = ycent + sqrt(dmax(0., rad * rad - xx * xx));
            YI(4, ix, iz) = +LARGEPOS;
          } else {
            YI(1, ix, iz) = -LARGEPOS;
            for (k = 2; k <= 4; k++) YI(k, ix, iz) = +LARGEPOS;
          }
        } else {
          for (k = 1; k <= 4; k++) YI(k, ix, iz) = LARGEPOS;
        }
      }
    }
  }
  for (iy = 0; iy <= ny; iy++) {
    double yy = ymmin + iy * dy - ycent;
    for (ix = 0; ix <= nx; ix++) {
      double xx = xmmin + ix * dx - xcent;
      double rsq = xx * xx + yy * yy;
      int in = outside ? (rsq >= rad * rad) : (rsq <= rad * rad);
      if (in) {
        ZI(1, ix, iy) = zcent + zmin;
        ZI(2, ix, iy) = zcent + zmax;
      } else {
        ZI(1, ix, iy) = LARGEPOS;
        ZI(2, ix, iy) = LARGEPOS;
      }
    }
  }
}

/* intserticpt, the insertion sort internal to ZSrfrvConductorFnew
 * (f3d_conductors.F:7803-7831).  v is the line's intercept list, 1-based. */
static void intserticpt(double ic, int64_t *nicpt, double *v) {
  int64_t ii;
  *nicpt = *nicpt + 1;
  if (*nicpt == 1) {
    v[0] = ic;
    return;
  }
  for (ii = *nicpt - 1; ii >= 1; ii--) {
    if (ic >= v[ii - 1]) {
      v[ii] = ic;
      break;
    } else {
      v[ii] = v[ii - 1];
    }
  }
  if (ic < v[0]) v[0] = ic;
}

/* Flat-segment clean-up for one line, f3d_conductors.F:7318-7397 (x),
 * 7519-7584 (y): a = lower, b = upper end of the flat segment in the
 * transverse coordinate, c the centre offset. */
static void flatcleanup_xy(double *v, int64_t *n, double a, double b, double cen) {
  int64_t ic = 1, k;
  while (ic < *n) {
    double *I = v - 1; /* 1-based view */
    if (I[ic] < -b + cen && I[ic + 1] == -b + cen) {
      I[ic + 1] = -a + cen;
    } else if (I[ic] == -a + cen && I[ic + 1] > -a + cen) {
      I[ic] = -b + cen;
      if (ic > 1) {
        if (I[ic - 1] == -a + cen) {
          for (k = ic - 1; k <= *n - 3; k++) I[k] = I[k + 2];
          *n = *n - 2;
          ic = ic - 2;
        }
      }
    } else if (I[ic] > -b + cen && I[ic + 1] < -a + cen) {
      for (k = ic; k <= *n - 2; k++) I[k] = I[k + 2];
      *n = *n - 2;
      ic = ic - 2;
    } else if (I[ic] < a + cen && I[ic + 1] == a + cen) {
      I[ic + 1] = b + cen;
    } else if (I[ic] == b + cen && I[ic + 1] > b + cen) {
      I[ic] = a + cen;
      if (ic > 1) {
        if (I[ic - 1] == b + cen) {
          for (k = ic - 1; k <= *n - 3; k++) I[k] = I[k + 2];
          *n = *n - 2;
          ic = ic - 2;
        }
      }
    } else if (I[ic] > a + cen && I[ic + 1] < b + cen) {
      for (k = ic; k <= *n - 2; k++) I[k] = I[k + 2];
      *n = *n - 2;
      ic = ic - 2;
    }
    ic = ic + 2;
  }
}

/* The shifts of the Fortran array sections above copy with overlapping
 * source and destination in increasing index order, which matches the
 * element-wise loops used here since the destination is below the source. */

/* ZSrfrvConductorFnew (f3d_conductors.F:7040-7855) restricted to straight
 * segments (rad = LARGEPOS for every segment), as used by ZAnnulus and other
 * polygonal ZSrfrvIn objects.  The caller has set nxicpt = nyicpt = nzicpt =
 * 2*(nn-1).  counts returns the largest number of intercepts used on any
 * line (for checking). */
void oc_zsrfrv_intercepts(int64_t nn, const double *rsrf, const double *zsrf,
                          double xcent, double ycent, double zcent,
                          oc_intercepts *icpt, int64_t count_only,
                          int64_t *counts) {
  double xmmin = icpt->xmmin, ymmin = icpt->ymmin, zmmin = icpt->zmmin;
  double dx = icpt->dx, dy = icpt->dy, dz = icpt->dz;
  int64_t nx = icpt->nx, ny = icpt->ny, nz = icpt->nz;
  int64_t ns = nn - 1, ii, ix, iy, iz, k;
  double *rmin = (double *)malloc(sizeof(double) * (ns + 1));
  double *rmax = (double *)malloc(sizeof(double) * (ns + 1));
  double *zmin = (double *)malloc(sizeof(double) * (ns + 1));
  double *zmax = (double *)malloc(sizeof(double) * (ns + 1));
  int64_t *flat = (int64_t *)malloc(sizeof(int64_t) * (2 * ns + 2));
  /* 1-based views of the segment data */
  const double *R = rsrf - 1, *Z = zsrf - 1;
  for (ii = 1; ii <= ns; ii++) {
    rmin[ii] = dmin(R[ii], R[ii + 1]);
    rmax[ii] = dmax(R[ii], R[ii + 1]);
    zmin[ii] = dmin(Z[ii], Z[ii + 1]);
    zmax[ii] = dmax(Z[ii], Z[ii + 1]);
  }
  int64_t maxn = 0;
  (void)count_only;
  /* --- x lines */
  for (iz = 0; iz <= nz; iz++) {
    double zz = zmmin + iz * dz - zcent;
    for (iy = 0; iy <= ny; iy++) {
      double yy = ymmin + iy * dy - ycent;
      double *v = &XI(1, iy, iz);
      int64_t nicpt = 0, nflat = 0;
      for (ii = 1; ii <= ns; ii++) {
        int ll = (zmin[ii] < zz && zz < zmax[ii]) || zz == Z[ii];
        if (zz == Z[ii + 1]) {
          int64_t iip1 = ii + 1;
          if (ii == nn - 1) iip1 = 1;
          if ((zmin[ii] < zz && zmin[iip1] < zz) || (zmax[ii] > zz && zmax[iip1] > zz)) ll = 1;
        }
        if (!ll) continue;
        if (Z[ii + 1] - Z[ii] == 0.) {
          flat[++nflat] = ii;
          int64_t iim1 = ii - 1;
          if (iim1 == 0) iim1 = nn - 1;
          int64_t iip1 = ii + 1;
          if (iip1 == nn) iip1 = 1;
          while (zmin[iip1] == zmax[iip1] && iip1 != ii) {
            iip1 = iip1 + 1;
            if (iip1 == nn) iip1 = 1;
          }
          if ((zmin[iim1] < zz && zmin[iip1] < zz) || (zmax[iim1] > zz && zmax[iip1] > zz)) {
            double rr = R[ii];
            if (yy * yy <= rr * rr) {
              double x1 = -sqrt(dmax(0., rr * rr - yy * yy));
              double x2 = +sqrt(dmax(0., rr * rr - yy * yy));
              intserticpt(x1 + xcent, &nicpt, v);
              intserticpt(x2 + xcent, &nicpt, v);
            }
          }
        } else {
          double ww = (zz - Z[ii]) / (Z[ii + 1] - Z[ii]);
          double rr = R[ii] * (1. - ww) + R[ii + 1] * ww;
          if (yy * yy <= rr * rr) {
            double x1 = -sqrt(dmax(0., rr * rr - yy * yy));
            double x2 = +sqrt(dmax(0., rr * rr - yy * yy));
            intserticpt(x1 + xcent, &nicpt, v);
            intserticpt(x2 + xcent, &nicpt, v);
          }
        }
      }
      for (k = 1; k <= nflat; k++) {
        ii = flat[k];
        if (yy > dmax(R[ii], R[ii + 1])) continue;
        double rlo = dmin(R[ii], R[ii + 1]), rhi = dmax(R[ii], R[ii + 1]);
        double a = sqrt(dmax(0., rlo * rlo - yy * yy));
        double b = sqrt(dmax(0., rhi * rhi - yy * yy));
        flatcleanup_xy(v, &nicpt, a, b, xcent);
      }
      if (nicpt > maxn) maxn = nicpt;
      for (k = nicpt + 1; k <= icpt->nxicpt; k++) v[k - 1] = LARGEPOS;
    }
  }
  /* --- y lines */
  for (iz = 0; iz <= nz; iz++) {
    double zz = zmmin + iz * dz - zcent;
    for (ix = 0; ix <= nx; ix++) {
      double xx = xmmin + ix * dx - xcent;
      double *v = &YI(1, ix, iz);
      int64_t nicpt = 0, nflat = 0;
      for (ii = 1; ii <= ns; ii++) {
        int ll = (zmin[ii] < zz && zz < zmax[ii]) || zz == Z[ii];
        if (zz == Z[ii + 1]) {
          int64_t iip1 = ii + 1;
          if (ii == nn - 1) iip1 = 1;
          if ((zmin[ii] < zz && zmin[iip1] < zz) || (zmax[ii] > zz && zmax[iip1] > zz)) ll = 1;
        }
        if (!ll) continue;
        if (Z[ii + 1] - Z[ii] == 0.) {
          flat[++nflat] = ii;
          int64_t iim1 = ii - 1;
          if (iim1 == 0) iim1 = nn - 1;
          int64_t iip1 = ii + 1;
          if (iip1 == nn) iip1 = 1;
          while (zmin[iip1] == zmax[iip1] && iip1 != ii) {
            iip1 = iip1 + 1;
            if (iip1 == nn) iip1 = 1;
          }
          if ((zmin[iim1] < zz && zmin[iip1] < zz) || (zmax[iim1] > zz && zmax[iip1] > zz)) {
            double rr = R[ii];
            if (xx * xx <= rr * rr) {
              double y1 = -sqrt(dmax(0., rr * rr - xx * xx));
              double y2 = +sqrt(dmax(0., rr * rr - xx * xx));
              intserticpt(y1 + ycent, &nicpt, v);
              intserticpt(y2 + ycent, &nicpt, v);
            }
          }
        } else {
          double ww = (zz - Z[ii]) / (Z[ii + 1] - Z[ii]);
          double rr = R[ii] * (1. - ww) + R[ii + 1] * ww;
          if (xx * xx <= rr * rr) {
            double y1 = -sqrt(dmax(0., rr * rr - xx * xx));
            double y2 = +sqrt(dmax(0., rr * rr - xx * xx));
            intserticpt(y1 + ycent, &nicpt, v);
            intserticpt(y2 + ycent, &nicpt, v);
          }
        }
      }
      for (k = 1; k <= nflat; k++) {
        ii = flat[k];
        if (xx > dmax(R[ii], R[ii + 1])) continue;
        double rlo = dmin(R[ii], R[ii + 1]), rhi = dmax(R[ii], R[ii + 1]);
        double a = sqrt(dmax(0., rlo * rlo - xx * xx));
        double b = sqrt(dmax(0., rhi * rhi - xx * xx));
        flatcleanup_xy(v, &nicpt, a, b, ycent);
      }
      if (nicpt > maxn) maxn = nicpt;
      for (k = nicpt + 1; k <= icpt->nyicpt; k++) v[k - 1] = LARGEPOS;
    }
  }
  /* --- z lines */
  for (iy = 0; iy <= ny; iy++) {
    double yy = ymmin + iy * dy - ycent;
    for (ix = 0; ix <= nx; ix++) {
      double xx = xmmin + ix * dx - xcent;
      double rr = sqrt(xx * xx + yy * yy);
      double *v = &ZI(1, ix, iy);
      int64_t nicpt = 0, nflat = 0;
      for (ii = 1; ii <= ns; ii++) {
        int ll = (rmin[ii] < rr && rr < rmax[ii]) || rr == R[ii];
        if (rr == R[ii + 1]) {
          int64_t iip1 = ii + 1;
          if (ii == nn - 1) iip1 = 1;
          if ((rmin[ii] < rr && rmin[iip1] < rr) || (rmax[ii] > rr && rmax[iip1] > rr)) ll = 1;
        }
        if (!ll) continue;
        if (R[ii + 1] - R[ii] == 0.) {
          flat[++nflat] = ii;
          int64_t iim1 = ii - 1;
          if (iim1 == 0) iim1 = nn - 1;
          int64_t iip1 = ii + 1;
          if (iip1 == nn) iip1 = 1;
          while (rmin[iip1] == rmax[iip1] && iip1 != ii) {
            iip1 = iip1 + 1;
            if (iip1 == nn) iip1 = 1;
          }
          if ((rmin[iim1] < rr && rmin[iip1] < rr) || (rmax[iim1] > rr && rmax[iip1] > rr)) {
            double zz = Z[ii];
            intserticpt(zz + zcent, &nicpt, v);
          }
        } else {
          double ww = (rr - R[ii]) / (R[ii + 1] - R[ii]);
          double zz = Z[ii] * (1. - ww) + Z[ii + 1] * ww;
          intserticpt(zz + zcent, &nicpt, v);
        }
      }
      for (k = 1; k <= nflat; k++) {
        ii = flat[k];
        double z1 = dmin(Z[ii], Z[ii + 1]), z2 = dmax(Z[ii], Z[ii + 1]);
        int64_t ic = 1, m;
        double *I = v - 1;
        while (ic < nicpt) {
          if (I[ic] < z1 + zcent && I[ic + 1] == z1 + zcent) {
            I[ic + 1] = z2 + zcent;
          } else if (I[ic] == z2 + zcent && I[ic + 1] > z2 + zcent) {
            I[ic] = z1 + zcent;
            if (ic > 1) {
              if (I[ic - 1] == z2 + zcent) {
                for (m = ic - 1; m <= nicpt - 3; m++) I[m] = I[m + 2];
                nicpt = nicpt - 2;
                ic = ic - 2;
              }
            }
          } else if (I[ic] > z1 + zcent && I[ic + 1] < z2 + zcent) {
            for (m = ic; m <= nicpt - 2; m++) I[m] = I[m + 2];
            nicpt = nicpt - 2;
            ic = ic - 2;
          }
          ic = ic + 2;
        }
      }
      if (nicpt > maxn) maxn = nicpt;
      for (k = nicpt + 1; k <= icpt->nzicpt; k++) v[k - 1] = LARGEPOS;
    }
  }
  if (counts) counts[0] = maxn;
  free(rmin); free(rmax); free(zmin); free(zmax); free(flat);
}

/* conductordelfromintercepts, f3d_conductors.F:12-728.  Appends the
 * interior and subgrid points to c (c->int_n, c->sub_n[] are the current
 * list lengths and must have room).  With count_only, only counts[0..2]
 * (interior, even, odd) are computed. */
int64_t oc_conductordelfromintercepts(const oc_intercepts *icpt,
                                      oc_conductors *c, double dfill,
                                      double fuzz, int64_t count_only,
                                      int64_t *counts) {
  if (icpt->nx < 0 || icpt->ny < 0 || icpt->nz < 0) {
    if (counts) counts[0] = counts[1] = counts[2] = 0;
    return 0;
  }
  double xmmin = icpt->xmmin, ymmin = icpt->ymmin, zmmin = icpt->zmmin;
  double dx = icpt->dx, dy = icpt->dy, dz = icpt->dz;
  int64_t nx = icpt->nx, ny = icpt->ny, nz = icpt->nz;
  int64_t ixl = icpt->ix, iyl = icpt->iy, izl = icpt->iz;
  int64_t nxicpt = icpt->nxicpt, nyicpt = icpt->nyicpt, nzicpt = icpt->nzicpt;
  double xmmax = xmmin + nx * dx, ymmax = ymmin + ny * dy, zmmax = zmmin + nz * dz;
  double dxi = 1. / dx, dyi = 1. / dy, dzi = 1. / dz;
  int64_t nc = 0, ne = 0, no = 0;
  int64_t ix, iy, iz, ii, k;
  int64_t ix1, ix2, iy1, iz1;
  double delx, dely, delz, xx, yy, zz, xside, yside, zside;
  double dels[6], volt[6];
  int64_t numb[6], indx[3];
  int64_t idel;

#define SAVE_SUBGRID()                                                        \
  do {                                                                        \
    int64_t par = (ix + ixl + iy + iyl + iz + izl) % 2;                       \
    if (count_only) {                                                         \
      if (par == 0) ne++; else no++;                                          \
    } else {                                                                  \
      int64_t p = par, n = c->sub_n[p];                                       \
      for (k = 0; k < 3; k++) c->sub_indx[p][3 * n + k] = indx[k];            \
      for (k = 0; k < 6; k++) {                                               \
        c->sub_dels[p][6 * n + k] = dels[k];                                  \
        c->sub_volt[p][6 * n + k] = volt[k];                                  \
        c->sub_numb[p][6 * n + k] = numb[k];                                  \
      }                                                                       \
      c->sub_ilevel[p][n] = icpt->mglevel;                                    \
      c->sub_n[p] = n + 1;                                                    \
      if (par == 0) ne++; else no++;                                          \
    }                                                                         \
  } while (0)

  /* --- Interior points, scanning x intercepts (f3d_conductors.F:180-267). */
  for (iz = 0; iz <= nz; iz++) {
    zz = zmmin + iz * dz;
    for (iy = 0; iy <= ny; iy++) {
      yy = ymmin + iy * dy;
      ii = 1;
      while (ii <= nxicpt) {
        if (XI(ii, iy, iz) >= LARGEPOS) break;
        double x1 = XI(ii, iy, iz);
        ii = ii + 1;
        double x2 = XI(ii, iy, iz);
        if (x2 >= xmmin && x1 <= xmmax) {
          ix1 = (int64_t)((dmax(x1 - fuzz, xmmin) - xmmin) * dxi);
          ix2 = (int64_t)((dmin(x2 + fuzz, xmmax) - xmmin) * dxi);
          if (xmmin + ix1 * dx <= x1 && ix1 < nx) ix1 = ix1 + 1;
          if (xmmin + (ix2 + 1) * dx <= x2 && ix2 < nx) ix2 = ix2 + 1;
          for (ix = ix1; ix <= ix2; ix++) {
            if (dfill < LARGEPOS) {
              xx = xmmin + ix * dx;
              delx = dmin(xx - x1, x2 - xx);
              dely = LARGEPOS;
              delz = LARGEPOS;
              for (iy1 = 1; iy1 <= nyicpt; iy1++) {
                dely = dmin(dely, fabs(yy - YI(iy1, ix, iz)));
                if (YI(iy1, ix, iz) > yy) break;
              }
              for (iz1 = 1; iz1 <= nzicpt; iz1++) {
                delz = dmin(delz, fabs(zz - ZI(iz1, ix, iy)));
                if (ZI(iz1, ix, iy) > zz) break;
              }
            } else {
              delx = 0.; dely = 0.; delz = 0.;
            }
            if (delx < dfill * dx || dely < dfill * dy || delz < dfill * dz) {
              if (!count_only) {
                int64_t n = c->int_n;
                c->int_indx[3 * n] = ix + ixl;
                c->int_indx[3 * n + 1] = iy + iyl;
                c->int_indx[3 * n + 2] = iz + izl;
                c->int_volt[n] = XV(ii / 2, iy, iz);
                c->int_numb[n] = XC(ii / 2, iy, iz);
                c->int_ilevel[n] = icpt->mglevel;
                c->int_n = n + 1;
              }
              nc++;
            }
          }
        }
        ii = ii + 1;
      }
    }
  }

  /* --- Subgrid points from the x intercepts (f3d_conductors.F:274-432). */
  for (iz = 0; iz <= nz; iz++) {
    zz = zmmin + iz * dz;
    for (iy = 0; iy <= ny; iy++) {
      yy = ymmin + iy * dy;
      ii = 0;
      xside = -1.;
      while (ii + 1 <= nxicpt) {
        if (XI(ii + 1, iy, iz) >= LARGEPOS) break;
        ii = ii + 1;
        xside = -xside;
        if (!(xmmin - dx < XI(ii, iy, iz) && XI(ii, iy, iz) < xmmax + dx)) continue;
        ix = (int64_t)floor((XI(ii, iy, iz) - xmmin) * dxi);
        if (xside > 0.) {
          if (ix < 0) continue;
          xx = xmmin + ix * dx;
          if (ii == 1) delx = (XI(ii, iy, iz) - xmmin) * dxi - ix;
          else if (xx > XI(ii - 1, iy, iz)) delx = (XI(ii, iy, iz) - xmmin) * dxi - ix;
          else delx = LARGEPOS;
          idel = 1;
        } else {
          ix = ix + 1;
          if (ix > nx) continue;
          xx = xmmin + ix * dx;
          if (ii == nxicpt) delx = ix - (XI(ii, iy, iz) - xmmin) * dxi;
          else if (xx < XI(ii + 1, iy, iz)) delx = ix - (XI(ii, iy, iz) - xmmin) * dxi;
          else delx = LARGEPOS;
          idel = 0;
        }
        if (0. < delx && delx < 1. - fuzz) {
          if (count_only) {
            if ((ix + ixl + iy + iyl + iz + izl) % 2 == 0) ne++; else no++;
            continue;
          }
          for (k = 0; k < 6; k++) { dels[k] = LARGEPOS; volt[k] = 0.; numb[k] = 0; }
          indx[0] = ix + ixl; indx[1] = iy + iyl; indx[2] = iz + izl;
          dels[idel] = delx;
          volt[idel] = XV((ii + 1) / 2, iy, iz);
          numb[idel] = XC((ii + 1) / 2, iy, iz);
          iy1 = 1;
          yside = +1.;
          while (iy1 <= nyicpt) {
            if (YI(iy1, ix, iz) >= LARGEPOS) break;
            if (yside > 0.) {
              if (iy1 == 1) dely = (YI(iy1, ix, iz) - yy) * dyi;
              else if (yy > YI(iy1 - 1, ix, iz)) dely = (YI(iy1, ix, iz) - yy) * dyi;
              else dely = LARGEPOS;
              idel = 3;
            } else {
              if (iy1 == nyicpt) dely = (yy - YI(iy1, ix, iz)) * dyi;
              else if (yy < YI(iy1 + 1, ix, iz)) dely = (yy - YI(iy1, ix, iz)) * dyi;
              else dely = LARGEPOS;
              idel = 2;
            }
            if (0. < dely && dely < 1. - fuzz && dely < dels[idel]) {
              dels[idel] = dely;
              volt[idel] = YV((iy1 + 1) / 2, ix, iz);
              numb[idel] = YC((iy1 + 1) / 2, ix, iz);
            }
            iy1 = iy1 + 1;
            yside = -yside;
          }
          iz1 = 1;
          zside = +1.;
          while (iz1 <= nzicpt) {
            if (ZI(iz1, ix, iy) >= LARGEPOS) break;
            if (zside > 0.) {
              if (iz1 == 1) delz = (ZI(iz1, ix, iy) - zz) * dzi;
              else if (zz > ZI(iz1 - 1, ix, iy)) delz = (ZI(iz1, ix, iy) - zz) * dzi;
              else delz = LARGEPOS;
              idel = 5;
            } else {
              if (iz1 == nzicpt) delz = (zz - ZI(iz1, ix, iy)) * dzi;
              else if (zz < ZI(iz1 + 1, ix, iy)) delz = (zz - ZI(iz1, ix, iy)) * dzi;
              else delz = LARGEPOS;
              idel = 4;
            }
            if (0. < delz && delz < 1. - fuzz && delz < dels[idel]) {
              dels[idel] = delz;
              volt[idel] = ZV((iz1 + 1) / 2, ix, iy);
              numb[idel] = ZC((iz1 + 1) / 2, ix, iy);
            }
            iz1 = iz1 + 1;
            zside = -zside;
          }
          SAVE_SUBGRID();
        }
      }
    }
  }

  /* --- Subgrid points from the y intercepts (f3d_conductors.F:434-584). */
  for (iz = 0; iz <= nz; iz++) {
    zz = zmmin + iz * dz;
    for (ix = 0; ix <= nx; ix++) {
      xx = xmmin + ix * dx;
      ii = 0;
      yside = -1.;
      while (ii + 1 <= nyicpt) {
        if (YI(ii + 1, ix, iz) >= LARGEPOS) break;
        ii = ii + 1;
        yside = -yside;
        if (!(ymmin - dy < YI(ii, ix, iz) && YI(ii, ix, iz) < ymmax + dy)) continue;
        iy = (int64_t)floor((YI(ii, ix, iz) - ymmin) * dyi);
        if (yside > 0.) {
          if (iy < 0) continue;
          yy = ymmin + iy * dy;
          if (ii == 1) dely = (YI(ii, ix, iz) - ymmin) * dyi - iy;
          else if (yy > YI(ii - 1, ix, iz)) dely = (YI(ii, ix, iz) - ymmin) * dyi - iy;
          else dely = LARGEPOS;
          idel = 3;
        } else {
          iy = iy + 1;
          if (iy > ny) continue;
          yy = ymmin + iy * dy;
          if (ii == nyicpt) dely = iy - (YI(ii, ix, iz) - ymmin) * dyi;
          else if (yy < YI(ii + 1, ix, iz)) dely = iy - (YI(ii, ix, iz) - ymmin) * dyi;
          else dely = LARGEPOS;
          idel = 2;
        }
        ix1 = 1;
        xside = +1.;
        delx = 2.;
        while (ix1 <= nxicpt) {
          if (XI(ix1, iy, iz) >= LARGEPOS) break;
          if (delx < 1.) break;
          if (xside > 0.) {
            if (XI(ix1, iy, iz) >= xx) delx = (XI(ix1, iy, iz) - xx) * dxi;
          } else {
            if (XI(ix1, iy, iz) <= xx) delx = (xx - XI(ix1, iy, iz)) * dxi;
          }
          ix1 = ix1 + 1;
          xside = -xside;
        }
        if (0. < dely && dely < 1. - fuzz && delx >= 1. - fuzz) {
          if (count_only) {
            if ((ix + ixl + iy + iyl + iz + izl) % 2 == 0) ne++; else no++;
            continue;
          }
          for (k = 0; k < 6; k++) { dels[k] = LARGEPOS; volt[k] = 0.; numb[k] = 0; }
          indx[0] = ix + ixl; indx[1] = iy + iyl; indx[2] = iz + izl;
          dels[idel] = dely;
          volt[idel] = YV((ii + 1) / 2, ix, iz);
          numb[idel] = YC((ii + 1) / 2, ix, iz);
          iz1 = 1;
          zside = +1.;
          while (iz1 <= nzicpt) {
            if (ZI(iz1, ix, iy) >= LARGEPOS) break;
            if (zside > 0.) {
              if (iz1 == 1) delz = (ZI(iz1, ix, iy) - zz) * dzi;
              else if (zz > ZI(iz1 - 1, ix, iy)) delz = (ZI(iz1, ix, iy) - zz) * dzi;
              else delz = LARGEPOS;
              idel = 5;
            } else {
              if (iz1 == nzicpt) delz = (zz - ZI(iz1, ix, iy)) * dzi;
              else if (zz < ZI(iz1 + 1, ix, iy)) delz = (zz - ZI(iz1, ix, iy)) * dzi;
              else delz = LARGEPOS;
              idel = 4;
            }
            if (0. < delz && delz < 1. - fuzz && delz < dels[idel]) {
              dels[idel] = delz;
              volt[idel] = ZV((iz1 + 1) / 2, ix, iy);
              numb[idel] = ZC((iz1 + 1) / 2, ix, iy);
            }
            iz1 = iz1 + 1;
            zside = -zside;
          }
          SAVE_SUBGRID();
        }
      }
    }
  }

  /* --- Subgrid points from the z intercepts (f3d_conductors.F:586-719). */
  for (iy = 0; iy <= ny; iy++) {
    yy = ymmin + iy * dy;
    for (ix = 0; ix <= nx; ix++) {
      xx = xmmin + ix * dx;
      ii = 0;
      zside = -1.;
      while (ii + 1 <= nzicpt) {
        if (ZI(ii + 1, ix, iy) >= LARGEPOS) break;
        ii = ii + 1;
        zside = -zside;
        if (!(zmmin - dz < ZI(ii, ix, iy) && ZI(ii, ix, iy) < zmmax + dz)) continue;
        iz = (int64_t)floor((ZI(ii, ix, iy) - zmmin) * dzi);
        if (zside > 0.) {
          if (iz < 0) continue;
          zz = zmmin + iz * dz;
          if (ii == 1) delz = (ZI(ii, ix, iy) - zmmin) * dzi - iz;
          else if (zz > ZI(ii - 1, ix, iy)) delz = (ZI(ii, ix, iy) - zmmin) * dzi - iz;
          else delz = LARGEPOS;
          idel = 5;
        } else {
          iz = iz + 1;
          if (iz > nz) continue;
          zz = zmmin + iz * dz;
          if (ii == nzicpt) delz = iz - (ZI(ii, ix, iy) - zmmin) * dzi;
          else if (zz < ZI(ii + 1, ix, iy)) delz = iz - (ZI(ii, ix, iy) - zmmin) * dzi;
          else delz = LARGEPOS;
          idel = 4;
        }
        ix1 = 1;
        xside = +1.;
        delx = 2.;
        while (ix1 <= nxicpt) {
          if (XI(ix1, iy, iz) >= LARGEPOS) break;
          if (delx < 1.) break;
          if (xside > 0.) {
            if (XI(ix1, iy, iz) >= xx) delx = (XI(ix1, iy, iz) - xx) * dxi;
          } else {
            if (XI(ix1, iy, iz) <= xx) delx = (xx - XI(ix1, iy, iz)) * dxi;
          }
          ix1 = ix1 + 1;
          xside = -xside;
        }
        iy1 = 1;
        yside = +1.;
        dely = 2.;
        while (iy1 <= nyicpt) {
          if (YI(iy1, ix, iz) >= LARGEPOS) break;
          if (dely < 1.) break;
          if (yside > 0.) {
            if (YI(iy1, ix, iz) >= yy) dely = (YI(iy1, ix, iz) - yy) * dyi;
          } else {
            if (YI(iy1, ix, iz) <= yy) dely = (yy - YI(iy1, ix, iz)) * dyi;
          }
          iy1 = iy1 + 1;
          yside = -yside;
        }
        if (0. < delz && delz < 1. - fuzz && delx >= 1. - fuzz && dely >= 1. - fuzz) {
          if (count_only) {
            if ((ix + ixl + iy + iyl + iz + izl) % 2 == 0) ne++; else no++;
            continue;
          }
          for (k = 0; k < 6; k++) { dels[k] = LARGEPOS; volt[k] = 0.; numb[k] = 0; }
          indx[0] = ix + ixl; indx[1] = iy + iyl; indx[2] = iz + izl;
          dels[idel] = delz;
          volt[idel] = ZV((ii + 1) / 2, ix, iy);
          numb[idel] = ZC((ii + 1) / 2, ix, iy);
          SAVE_SUBGRID();
        }
      }
    }
  }
#undef SAVE_SUBGRID
  if (counts) { counts[0] = nc; counts[1] = ne; counts[2] = no; }
  return 0;
}
