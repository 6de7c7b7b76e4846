/*
 * oc_correctede.c -- TEST INFRASTRUCTURE ONLY (parity oracle).
 *
 * FP64 restatement of the reference's near-conductor E-field correction
 * (f3d.lcorrectede) from pywarp90/f3d_mgrid.F: the node classification
 * icgrid (setupconductorfielddata 4581-4633), the one-sided E at subgrid
 * points (getefieldatconductorsubgrid 4635-4735), the doubling of the grid
 * E at interior conductor points (fixefieldatconductorpoints 4737-4772) and
 * the particle gather that uses them (sete3dwithconductor 4774-5010 with
 * GETEEFROME 5182-5233), efetch 1 and 3, no symmetries, level 0 only,
 * nxguardphi = 1, nxguarde = 0.
 *
 * icgrid(0:nx, 0:ny, 0:nz): 0 vacuum, OC_INTERIOR interior conductor point,
 * +(i+1) even subgrid record i, -(i+1) odd subgrid record i.  The
 * reference's sentinel nmax+1 only has to exceed every record index.
 * Reads at index n+1 (a particle exactly on an upper face, whose weight
 * there is zero) take icgrid = 0 and the nearest in-range field value.
 */
#include "warp_oracle.h"

#define OC_INTERIOR 2147483647
#define IDXP(ix, iy, iz) \
  ((int64_t)((ix) + 1) + (nx + 3) * ((int64_t)((iy) + 1) + (ny + 3) * (int64_t)((iz) + 1)))
#define IDXR(ix, iy, iz) \
  ((int64_t)(ix) + (nx + 1) * ((int64_t)(iy) + (ny + 1) * (int64_t)(iz)))

static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* setupconductorfielddata: interior points first, then the even list,
 * then the odd list; later writes win. */
void oc_setup_icgrid(int64_t nx, int64_t ny, int64_t nz, int64_t nint,
                     const int64_t *int_indx, int64_t ne, const int64_t *e_indx,
                     int64_t no, const int64_t *o_indx, int32_t *icgrid) {
  int64_t i, n = (nx + 1) * (ny + 1) * (nz + 1);
  for (i = 0; i < n; i++) icgrid[i] = 0;
  for (i = 0; i < nint; i++)
    icgrid[IDXR(int_indx[3 * i], int_indx[3 * i + 1], int_indx[3 * i + 2])] = OC_INTERIOR;
  for (i = 0; i < ne; i++)
    icgrid[IDXR(e_indx[3 * i], e_indx[3 * i + 1], e_indx[3 * i + 2])] = (int32_t)(i + 1);
  for (i = 0; i < no; i++)
    icgrid[IDXR(o_indx[3 * i], o_indx[3 * i + 1], o_indx[3 * i + 2])] = (int32_t)(-(i + 1));
}

/* getefieldatconductorsubgrid for one list (SETEE for the three axes):
 * efield0(3, n) and efieldd(6, n) (efieldd entries are written only where
 * the neighbour on that side is an interior point). */
void oc_efield_subgrid(const double *phi, int64_t nx, int64_t ny, int64_t nz, double dx,
                       double dy, double dz, const int32_t *icgrid, int64_t n,
                       const int64_t *indx, const double *dels, const double *volt,
                       double *efield0, double *efieldd) {
  int64_t ic;
  for (ic = 0; ic < n; ic++) {
    int64_t ix = indx[3 * ic], iy = indx[3 * ic + 1], iz = indx[3 * ic + 2];
    int id;
    for (id = 0; id < 3; id++) {
      int64_t mx = id == 0 ? -1 : 0, my = id == 1 ? -1 : 0, mz = id == 2 ? -1 : 0;
      int64_t px = -mx, py = -my, pz = -mz;
      double di = id == 0 ? 1. / dx : (id == 1 ? 1. / dy : 1. / dz);
      int im = 2 * id, ip = 2 * id + 1;
      double vv0, dd0, vv1, dd1;
      const double *d = dels + 6 * ic, *v = volt + 6 * ic;
      if (0. < d[im] && d[im] <= 1.) {
        vv0 = v[im];
        dd0 = d[im];
      } else {
        vv0 = phi[IDXP(ix + mx, iy + my, iz + mz)];
        dd0 = 1.;
      }
      if (0. < d[ip] && d[ip] <= 1.) {
        vv1 = v[ip];
        dd1 = d[ip];
      } else {
        vv1 = phi[IDXP(ix + px, iy + py, iz + pz)];
        dd1 = 1.;
      }
      double ppp = phi[IDXP(ix, iy, iz)];
      efield0[3 * ic + id] =
          (dd1 / (dd0 + dd1) * (vv0 - ppp) / dd0 + dd0 / (dd0 + dd1) * (ppp - vv1) / dd1) * di;
      if (ix + mx >= 0 && iy + my >= 0 && iz + mz >= 0)
        if (icgrid[IDXR(ix + mx, iy + my, iz + mz)] == OC_INTERIOR)
          efieldd[6 * ic + im] = ((vv0 - ppp) / dd0 - (2. - dd0) / (2. + dd1) * (ppp - vv1) / dd1) *
                                 ((2. + dd1) / (dd1 + dd0)) * di;
      if (ix + px <= nx && iy + py <= ny && iz + pz <= nz)
        if (icgrid[IDXR(ix + px, iy + py, iz + pz)] == OC_INTERIOR)
          efieldd[6 * ic + ip] = ((ppp - vv1) / dd1 - (2. - dd1) / (2. + dd0) * (vv0 - ppp) / dd0) *
                                 ((2. + dd0) / (dd0 + dd1)) * di;
    }
  }
}

/* fixefieldatconductorpoints: selfe doubled at the interior points. */
void oc_fixefield(double *selfe, int64_t nx, int64_t ny, int64_t nz, int64_t nint,
                  const int64_t *int_indx) {
  int64_t i;
  for (i = 0; i < nint; i++) {
    double *e = selfe + 3 * IDXR(int_indx[3 * i], int_indx[3 * i + 1], int_indx[3 * i + 2]);
    e[0] = 2. * e[0];
    e[1] = 2. * e[1];
    e[2] = 2. * e[2];
  }
}

/* sete3dwithconductor, efetch 1 or 3: E added to ex, ey, ez. */
void oc_sete3d_conductor(const int32_t *icgrid, const double *e0_even, const double *ed_even,
                         const double *e0_odd, const double *ed_odd, const double *phi,
                         const double *selfe, int64_t np, const double *xp, const double *yp,
                         const double *zp, double zgrid, double xmmin, double ymmin,
                         double zmmin, double dx, double dy, double dz, int64_t nx, int64_t ny,
                         int64_t nz, int64_t efetch, double *ex, double *ey, double *ez) {
  double tdxi = 1. / (2. * dx), tdyi = 1. / (2. * dy), tdzi = 1. / (2. * dz);
  double dxi = 1. / dx, dyi = 1. / dy, dzi = 1. / dz;
  int64_t ip;
#define ICG(a, b, c) (((a) > nx || (b) > ny || (c) > nz) ? 0 : icgrid[IDXR(a, b, c)])
#define PHI(a, b, c) phi[IDXP(clampi(a, -1, nx + 1), clampi(b, -1, ny + 1), clampi(c, -1, nz + 1))]
#define SE(id, a, b, c) selfe[3 * IDXR(clampi(a, 0, nx), clampi(b, 0, ny), clampi(c, 0, nz)) + (id)]
  OC_PARFOR()
  for (ip = 0; ip < np; ip++) {
    double xi = (xp[ip] - xmmin) * dxi, yj = (yp[ip] - ymmin) * dyi;
    double zk = (zp[ip] - zgrid - zmmin) * dzi;
    if (!(xi >= 0. && xi <= nx && yj >= 0. && yj <= ny && zk >= 0. && zk <= nz)) continue;
    int64_t i = (int64_t)xi, j = (int64_t)yj, k = (int64_t)zk;
    double u1 = xi - i, v1 = yj - j, w1 = zk - k;
    double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
    double ext, eyt, ezt;
    if (ICG(i, j, k) == 0 && ICG(i + 1, j, k) == 0 && ICG(i, j + 1, k) == 0 &&
        ICG(i + 1, j + 1, k) == 0 && ICG(i, j, k + 1) == 0 && ICG(i + 1, j, k + 1) == 0 &&
        ICG(i, j + 1, k + 1) == 0 && ICG(i + 1, j + 1, k + 1) == 0) {
      if (efetch == 3) {
        double out[3];
        int c;
        for (c = 0; c < 3; c++)
          out[c] = u0 * v0 * w0 * SE(c, i, j, k) + u1 * v0 * w0 * SE(c, i + 1, j, k) +
                   u0 * v1 * w0 * SE(c, i, j + 1, k) + u1 * v1 * w0 * SE(c, i + 1, j + 1, k) +
                   u0 * v0 * w1 * SE(c, i, j, k + 1) + u1 * v0 * w1 * SE(c, i + 1, j, k + 1) +
                   u0 * v1 * w1 * SE(c, i, j + 1, k + 1) + u1 * v1 * w1 * SE(c, i + 1, j + 1, k + 1);
        ext = out[0];
        eyt = out[1];
        ezt = out[2];
      } else {
        ext = u0 * v0 * w0 * (PHI(i - 1, j, k) - PHI(i + 1, j, k)) * tdxi +
              u1 * v0 * w0 * (PHI(i, j, k) - PHI(i + 2, j, k)) * tdxi +
              u0 * v1 * w0 * (PHI(i - 1, j + 1, k) - PHI(i + 1, j + 1, k)) * tdxi +
              u1 * v1 * w0 * (PHI(i, j + 1, k) - PHI(i + 2, j + 1, k)) * tdxi +
              u0 * v0 * w1 * (PHI(i - 1, j, k + 1) - PHI(i + 1, j, k + 1)) * tdxi +
              u1 * v0 * w1 * (PHI(i, j, k + 1) - PHI(i + 2, j, k + 1)) * tdxi +
              u0 * v1 * w1 * (PHI(i - 1, j + 1, k + 1) - PHI(i + 1, j + 1, k + 1)) * tdxi +
              u1 * v1 * w1 * (PHI(i, j + 1, k + 1) - PHI(i + 2, j + 1, k + 1)) * tdxi;
        eyt = u0 * v0 * w0 * (PHI(i, j - 1, k) - PHI(i, j + 1, k)) * tdyi +
              u1 * v0 * w0 * (PHI(i + 1, j - 1, k) - PHI(i + 1, j + 1, k)) * tdyi +
              u0 * v1 * w0 * (PHI(i, j, k) - PHI(i, j + 2, k)) * tdyi +
              u1 * v1 * w0 * (PHI(i + 1, j, k) - PHI(i + 1, j + 2, k)) * tdyi +
              u0 * v0 * w1 * (PHI(i, j - 1, k + 1) - PHI(i, j + 1, k + 1)) * tdyi +
              u1 * v0 * w1 * (PHI(i + 1, j - 1, k + 1) - PHI(i + 1, j + 1, k + 1)) * tdyi +
              u0 * v1 * w1 * (PHI(i, j, k + 1) - PHI(i, j + 2, k + 1)) * tdyi +
              u1 * v1 * w1 * (PHI(i + 1, j, k + 1) - PHI(i + 1, j + 2, k + 1)) * tdyi;
        ezt = u0 * v0 * w0 * (PHI(i, j, k - 1) - PHI(i, j, k + 1)) * tdzi +
              u1 * v0 * w0 * (PHI(i + 1, j, k - 1) - PHI(i + 1, j, k + 1)) * tdzi +
              u0 * v1 * w0 * (PHI(i, j + 1, k - 1) - PHI(i, j + 1, k + 1)) * tdzi +
              u1 * v1 * w0 * (PHI(i + 1, j + 1, k - 1) - PHI(i + 1, j + 1, k + 1)) * tdzi +
              u0 * v0 * w1 * (PHI(i, j, k) - PHI(i, j, k + 2)) * tdzi +
              u1 * v0 * w1 * (PHI(i + 1, j, k) - PHI(i + 1, j, k + 2)) * tdzi +
              u0 * v1 * w1 * (PHI(i, j + 1, k) - PHI(i, j + 1, k + 2)) * tdzi +
              u1 * v1 * w1 * (PHI(i + 1, j + 1, k) - PHI(i + 1, j + 1, k + 2)) * tdzi;
      }
    } else {
      double ee[3] = {0., 0., 0.}, ww[3] = {0., 0., 0.};
      /* GETEEFROME(id, ox,oy,oz, px,py,pz, di, wm, wp) */
#define GETEE(id, ox, oy, oz, px, py, pz, di, wm, wp)                                           \
  do {                                                                                         \
    int32_t c0 = ICG(i + (ox), j + (oy), k + (oz));                                            \
    int32_t c1 = ICG(i + (ox) + (px), j + (oy) + (py), k + (oz) + (pz));                       \
    if (c0 == 0) {                                                                             \
      if (efetch == 3)                                                                         \
        ee[id] = ee[id] + (wm) * SE(id, i + (ox), j + (oy), k + (oz));                         \
      else                                                                                     \
        ee[id] = ee[id] + (wm) * (PHI(i + (ox) - (px), j + (oy) - (py), k + (oz) - (pz)) -     \
                                  PHI(i + (ox) + (px), j + (oy) + (py), k + (oz) + (pz))) *    \
                              0.5 * (di);                                                      \
      ww[id] = ww[id] + (wm);                                                                  \
    } else if (c0 != OC_INTERIOR) {                                                            \
      const double *e0 = c0 < 0 ? e0_odd + 3 * (-c0 - 1) : e0_even + 3 * (c0 - 1);            \
      const double *ed = c0 < 0 ? ed_odd + 6 * (-c0 - 1) : ed_even + 6 * (c0 - 1);             \
      ee[id] = ee[id] + (wm) * e0[id];                                                         \
      ww[id] = ww[id] + (wm);                                                                  \
      if (c1 == OC_INTERIOR) {                                                                 \
        ee[id] = ee[id] + (wp) * ed[2 * (id) + 1];                                             \
        ww[id] = ww[id] + (wp);                                                                \
      }                                                                                        \
    }                                                                                          \
    if (c1 == 0) {                                                                             \
      if (efetch == 3)                                                                         \
        ee[id] = ee[id] + (wp) * SE(id, i + (ox) + (px), j + (oy) + (py), k + (oz) + (pz));    \
      else                                                                                     \
        ee[id] = ee[id] + (wp) * (PHI(i + (ox), j + (oy), k + (oz)) -                          \
                                  PHI(i + (ox) + 2 * (px), j + (oy) + 2 * (py),                \
                                      k + (oz) + 2 * (pz))) *                                  \
                              0.5 * (di);                                                      \
      ww[id] = ww[id] + (wp);                                                                  \
    } else if (c1 != OC_INTERIOR) {                                                            \
      const double *e0 = c1 < 0 ? e0_odd + 3 * (-c1 - 1) : e0_even + 3 * (c1 - 1);            \
      const double *ed = c1 < 0 ? ed_odd + 6 * (-c1 - 1) : ed_even + 6 * (c1 - 1);             \
      ee[id] = ee[id] + (wp) * e0[id];                                                         \
      ww[id] = ww[id] + (wp);                                                                  \
      if (c0 == OC_INTERIOR) {                                                                 \
        ee[id] = ee[id] + (wm) * ed[2 * (id)];                                                 \
        ww[id] = ww[id] + (wm);                                                                \
      }                                                                                        \
    }                                                                                          \
  } while (0)
      GETEE(0, 0, 0, 0, 1, 0, 0, dxi, u0 * v0 * w0, u1 * v0 * w0);
      GETEE(0, 0, 1, 0, 1, 0, 0, dxi, u0 * v1 * w0, u1 * v1 * w0);
      GETEE(0, 0, 0, 1, 1, 0, 0, dxi, u0 * v0 * w1, u1 * v0 * w1);
      GETEE(0, 0, 1, 1, 1, 0, 0, dxi, u0 * v1 * w1, u1 * v1 * w1);
      if (ww[0] != 0.) ww[0] = 1. / ww[0];
      GETEE(1, 0, 0, 0, 0, 1, 0, dyi, u0 * v0 * w0, u0 * v1 * w0);
      GETEE(1, 1, 0, 0, 0, 1, 0, dyi, u1 * v0 * w0, u1 * v1 * w0);
      GETEE(1, 0, 0, 1, 0, 1, 0, dyi, u0 * v0 * w1, u0 * v1 * w1);
      GETEE(1, 1, 0, 1, 0, 1, 0, dyi, u1 * v0 * w1, u1 * v1 * w1);
      if (ww[1] != 0.) ww[1] = 1. / ww[1];
      GETEE(2, 0, 0, 0, 0, 0, 1, dzi, u0 * v0 * w0, u0 * v0 * w1);
      GETEE(2, 1, 0, 0, 0, 0, 1, dzi, u1 * v0 * w0, u1 * v0 * w1);
      GETEE(2, 0, 1, 0, 0, 0, 1, dzi, u0 * v1 * w0, u0 * v1 * w1);
      GETEE(2, 1, 1, 0, 0, 0, 1, dzi, u1 * v1 * w0, u1 * v1 * w1);
      if (ww[2] != 0.) ww[2] = 1. / ww[2];
#undef GETEE
      ext = ee[0] * ww[0];
      eyt = ee[1] * ww[1];
      ezt = ee[2] * ww[2];
    }
    ex[ip] = ex[ip] + ext;
    ey[ip] = ey[ip] + eyt;
    ez[ip] = ez[ip] + ezt;
  }
#undef ICG
#undef PHI
#undef SE
}
