/*
 * oc_mgrid.c -- TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline).
 *
 * FP64 restatement of the reference's 3-D multigrid Poisson solver with
 * subgrid-resolved conductors, pywarp90/f3d_mgrid.F, serial build
 * (no MPIPARALLEL, no bends, lprecalccoeffs=.false., nxguardphi=1,
 * nxguardrho=0, mgcoarsening=2).  Loop order and floating-point expression
 * order follow the Fortran so that results agree bit for bit with a
 * non-FMA x86 build of the reference.  Compile with -ffp-contract=off.
 */
#include "warp_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define MGCOARSENING 2

/* Index into an array dimensioned (-g:n+g) along each axis, x fastest. */
#define IDX(ix, iy, iz, nx, ny, g) \
  ((int64_t)((ix) + (g)) + ((int64_t)(nx) + 1 + 2 * (g)) * \
   ((int64_t)((iy) + (g)) + ((int64_t)(ny) + 1 + 2 * (g)) * (int64_t)((iz) + (g))))
#define PHI(ix, iy, iz) phi[IDX(ix, iy, iz, nx, ny, 1)]
#define RHO(ix, iy, iz) rho[IDX(ix, iy, iz, nx, ny, 0)]
#define RES(ix, iy, iz) res[IDX(ix, iy, iz, nx, ny, 1)]

static double dmin(double a, double b) { return a < b ? a : b; }

/* top.eps0 = 1/(mu0*clight*clight), mu0 = 4*pi/10000000 (scripts/warp.py:198-202,
 * pi and clight from pywarp90/top.v:172-175). */
#ifdef _OPENMP
#include <omp.h>
#endif

void oc_set_threads(int64_t n) {
#ifdef _OPENMP
  omp_set_num_threads((int)(n > 0 ? n : 1));
#else
  (void)n;
#endif
}

int64_t oc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

double oc_eps0(void) {
  double pi = 3.14159265358979323;
  double clight = 2.99792458e+8;
  double mu0 = 4 * pi / 10000000;
  return 1 / (mu0 * clight * clight);
}

/* applyboundaryconditions3d, f3d_mgrid.F:3847-3955 (ncomp=1). */
void oc_applyboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz,
                                  int64_t gx, int64_t gy, int64_t gz,
                                  double *u, const int64_t *bounds,
                                  int64_t lwithdirichlet, int64_t lzerodirichlet) {
  int64_t sx = nx + 1 + 2 * gx, sy = ny + 1 + 2 * gy;
#define U(i, j, k) u[((i) + gx) + sx * (((j) + gy) + sy * ((k) + gz))]
  int64_t ix, iy, iz, k;
  if (gx > 0) {
    if (lwithdirichlet) {
      if (bounds[0] == 0)
        for (ix = -1; ix >= -gx; ix--)
          for (iz = 0; iz <= nz; iz++)
            for (iy = 0; iy <= ny; iy++)
              U(ix, iy, iz) = 2. * U(ix + 1, iy, iz) - U(ix + 2, iy, iz);
      if (bounds[1] == 0)
        for (ix = nx + 1; ix <= nx + gx; ix++)
          for (iz = 0; iz <= nz; iz++)
            for (iy = 0; iy <= ny; iy++)
              U(ix, iy, iz) = 2. * U(ix - 1, iy, iz) - U(ix - 2, iy, iz);
    } else if (lzerodirichlet) {
      if (bounds[0] == 0)
        for (iz = -gz; iz <= nz + gz; iz++)
          for (iy = -gy; iy <= ny + gy; iy++)
            for (ix = -gx; ix <= -1; ix++) U(ix, iy, iz) = 0.;
      if (bounds[1] == 0)
        for (iz = -gz; iz <= nz + gz; iz++)
          for (iy = -gy; iy <= ny + gy; iy++)
            for (ix = nx + 1; ix <= nx + gx; ix++) U(ix, iy, iz) = 0.;
    }
    if (bounds[0] == 1)
      for (iz = 0; iz <= nz; iz++)
        for (iy = 0; iy <= ny; iy++)
          for (k = 1; k <= gx; k++) U(-k, iy, iz) = U(k, iy, iz);
    if (bounds[1] == 1)
      for (iz = 0; iz <= nz; iz++)
        for (iy = 0; iy <= ny; iy++)
          for (k = 1; k <= gx; k++) U(nx + k, iy, iz) = U(nx - k, iy, iz);
    if (bounds[0] == 2 && bounds[1] == 2) {
      for (iz = 0; iz <= nz; iz++)
        for (iy = 0; iy <= ny; iy++) {
          for (k = 1; k <= gx; k++) U(-k, iy, iz) = U(nx - k, iy, iz);
          for (k = 0; k <= gx; k++) U(nx + k, iy, iz) = U(k, iy, iz);
        }
    }
  }
  if (gy > 0) {
    if (lwithdirichlet) {
      if (bounds[2] == 0)
        for (iy = -1; iy >= -gy; iy--)
          for (iz = 0; iz <= nz; iz++)
            for (ix = -gx; ix <= nx + gx; ix++)
              U(ix, iy, iz) = 2. * U(ix, iy + 1, iz) - U(ix, iy + 2, iz);
      if (bounds[3] == 0)
        for (iy = ny + 1; iy <= ny + gy; iy++)
          for (iz = 0; iz <= nz; iz++)
            for (ix = -gx; ix <= nx + gx; ix++)
              U(ix, iy, iz) = 2. * U(ix, iy - 1, iz) - U(ix, iy - 2, iz);
    } else if (lzerodirichlet) {
      if (bounds[2] == 0)
        for (iz = -gz; iz <= nz + gz; iz++)
          for (iy = -gy; iy <= -1; iy++)
            for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, iz) = 0.;
      if (bounds[3] == 0)
        for (iz = -gz; iz <= nz + gz; iz++)
          for (iy = ny + 1; iy <= ny + gy; iy++)
            for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, iz) = 0.;
    }
    if (bounds[2] == 1)
      for (iz = 0; iz <= nz; iz++)
        for (k = 1; k <= gy; k++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, -k, iz) = U(ix, k, iz);
    if (bounds[3] == 1)
      for (iz = 0; iz <= nz; iz++)
        for (k = 1; k <= gy; k++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, ny + k, iz) = U(ix, ny - k, iz);
    if (bounds[2] == 2 && bounds[3] == 2)
      for (iz = 0; iz <= nz; iz++) {
        for (k = 1; k <= gy; k++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, -k, iz) = U(ix, ny - k, iz);
        for (k = 0; k <= gy; k++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, ny + k, iz) = U(ix, k, iz);
      }
  }
  if (gz > 0) {
    if (lwithdirichlet) {
      if (bounds[4] == 0)
        for (iz = -1; iz >= -gz; iz--)
          for (iy = -gy; iy <= ny + gy; iy++)
            for (ix = -gx; ix <= nx + gx; ix++)
              U(ix, iy, iz) = 2. * U(ix, iy, iz + 1) - U(ix, iy, iz + 2);
      if (bounds[5] == 0)
        for (iz = nz + 1; iz <= nz + gz; iz++)
          for (iy = -gy; iy <= ny + gy; iy++)
            for (ix = -gx; ix <= nx + gx; ix++)
              U(ix, iy, iz) = 2. * U(ix, iy, iz - 1) - U(ix, iy, iz - 2);
    } else if (lzerodirichlet) {
      if (bounds[4] == 0)
        for (iz = -gz; iz <= -1; iz++)
          for (iy = -gy; iy <= ny + gy; iy++)
            for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, iz) = 0.;
      if (bounds[5] == 0)
        for (iz = nz + 1; iz <= nz + gz; iz++)
          for (iy = -gy; iy <= ny + gy; iy++)
            for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, iz) = 0.;
    }
    if (bounds[4] == 1)
      for (k = 1; k <= gz; k++)
        for (iy = -gy; iy <= ny + gy; iy++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, -k) = U(ix, iy, k);
    if (bounds[5] == 1)
      for (k = 1; k <= gz; k++)
        for (iy = -gy; iy <= ny + gy; iy++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, nz + k) = U(ix, iy, nz - k);
    if (bounds[4] == 2 && bounds[5] == 2) {
      for (k = 1; k <= gz; k++)
        for (iy = -gy; iy <= ny + gy; iy++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, -k) = U(ix, iy, nz - k);
      for (k = 0; k <= gz; k++)
        for (iy = -gy; iy <= ny + gy; iy++)
          for (ix = -gx; ix <= nx + gx; ix++) U(ix, iy, nz + k) = U(ix, iy, k);
    }
  }
#undef U
}

/* getnextcoarselevel3d, f3d_mgrid.F:3601-3652 (mgcoarsening = 2). */
void oc_getnextcoarselevel3d(int64_t nx, int64_t ny, int64_t nz,
                             double dx, double dy, double dz,
                             int64_t *nxc, int64_t *nyc, int64_t *nzc,
                             double *dxc, double *dyc, double *dzc) {
  int64_t nxcoarse = (nx + 1) / MGCOARSENING;
  if (nxcoarse % 2 == 1) nxcoarse = nxcoarse + 1;
  int64_t nycoarse = (ny + 1) / MGCOARSENING;
  if (nycoarse % 2 == 1) nycoarse = nycoarse + 1;
  int64_t nzcoarse = (nz + 1) / MGCOARSENING;
  if (nzcoarse % 2 == 1) nzcoarse = nzcoarse + 1;
  double mingridsize = dmin(dx, dy);
  mingridsize = dmin(mingridsize, dz);
  if (dx > 4. / 3. * mingridsize) nxcoarse = nx;
  if (dy > 4. / 3. * mingridsize) nycoarse = ny;
  if (dz > 4. / 3. * mingridsize) nzcoarse = nz;
  *dxc = dx * nx / nxcoarse;
  if (nycoarse == 0) *dyc = *dxc;
  else *dyc = dy * ny / nycoarse;
  *dzc = dz * nz / nzcoarse;
  *nxc = nxcoarse;
  *nyc = nycoarse;
  *nzc = nzcoarse;
}

/* getmglevels, f3d_mgrid.F:3654-3845 (serial). */
void oc_getmglevels(int64_t nx, int64_t ny, int64_t nz,
                    double dx, double dy, double dz, oc_conductors *c) {
  int64_t l;
  for (l = 0; l < c->levels; l++)
    if (c->levellx[l] > 1 || c->levelly[l] > 1 || c->levellz[l] > 1) return;
  int64_t mglevel = 0;
  int64_t nxf = nx, nyf = ny, nzf = nz;
  double dxf = dx, dyf = dy, dzf = dz, lx = 1., ly = 1., lz = 1.;
  for (;;) {
    c->levelnx[mglevel] = nxf;
    c->levelny[mglevel] = nyf;
    c->levelnz[mglevel] = nzf;
    c->levellx[mglevel] = lx;
    c->levelly[mglevel] = ly;
    c->levellz[mglevel] = lz;
    c->levels = mglevel + 1;
    if (nxf < 4 || (nyf < 4 && nyf > 0) || nzf < 4 || mglevel == 100) break;
    int64_t nxc, nyc, nzc;
    double dxc, dyc, dzc;
    oc_getnextcoarselevel3d(nxf, nyf, nzf, dxf, dyf, dzf, &nxc, &nyc, &nzc,
                            &dxc, &dyc, &dzc);
    double cxf = dxc / dxf, cyf = dyc / dyf, czf = dzc / dzf;
    nxf = nxc; nyf = nyc; nzf = nzc;
    dxf = dxc; dyf = dyc; dzf = dzc;
    mglevel = mglevel + 1;
    lx = lx * cxf; ly = ly * cyf; lz = lz * czf;
  }
}

/* isortconductor + iswap/rswapconductor, f3d_mgrid.F:4078-4139: stable
 * partition of the list by level.  Points with a level outside
 * [0, levels) are dropped. */
static void sort_interior(oc_conductors *c) {
  int64_t nn = c->int_n, levels = c->levels, ii = 0, l, ic, k;
  int64_t *isort = (int64_t *)malloc(sizeof(int64_t) * (nn > 0 ? nn : 1));
  for (l = 0; l < levels; l++) {
    c->int_istart[l] = ii;
    for (ic = 0; ic < nn; ic++)
      if (c->int_ilevel[ic] == l) isort[ii++] = ic;
  }
  c->int_istart[levels] = ii;
  int64_t na = ii;
  int64_t *ti = (int64_t *)malloc(sizeof(int64_t) * 3 * (nn > 0 ? nn : 1));
  double *td = (double *)malloc(sizeof(double) * (nn > 0 ? nn : 1));
  memcpy(ti, c->int_indx, sizeof(int64_t) * 3 * nn);
  for (ic = 0; ic < na; ic++)
    for (k = 0; k < 3; k++) c->int_indx[3 * ic + k] = ti[3 * isort[ic] + k];
  memcpy(ti, c->int_numb, sizeof(int64_t) * nn);
  for (ic = 0; ic < na; ic++) c->int_numb[ic] = ti[isort[ic]];
  memcpy(td, c->int_volt, sizeof(double) * nn);
  for (ic = 0; ic < na; ic++) c->int_volt[ic] = td[isort[ic]];
  memcpy(ti, c->int_ilevel, sizeof(int64_t) * nn);
  for (ic = 0; ic < na; ic++) c->int_ilevel[ic] = ti[isort[ic]];
  c->int_n = na;
  free(isort); free(ti); free(td);
}

static void sort_subgrid(oc_conductors *c, int p) {
  int64_t nn = c->sub_n[p], levels = c->levels, ii = 0, l, ic, k;
  int64_t nn1 = nn > 0 ? nn : 1;
  int64_t *isort = (int64_t *)malloc(sizeof(int64_t) * nn1);
  for (l = 0; l < levels; l++) {
    c->sub_istart[p][l] = ii;
    for (ic = 0; ic < nn; ic++)
      if (c->sub_ilevel[p][ic] == l) isort[ii++] = ic;
  }
  c->sub_istart[p][levels] = ii;
  int64_t na = ii;
  int64_t *ti = (int64_t *)malloc(sizeof(int64_t) * 6 * nn1);
  double *td = (double *)malloc(sizeof(double) * 6 * nn1);
  memcpy(ti, c->sub_indx[p], sizeof(int64_t) * 3 * nn);
  for (ic = 0; ic < na; ic++)
    for (k = 0; k < 3; k++) c->sub_indx[p][3 * ic + k] = ti[3 * isort[ic] + k];
  memcpy(td, c->sub_dels[p], sizeof(double) * 6 * nn);
  for (ic = 0; ic < na; ic++)
    for (k = 0; k < 6; k++) c->sub_dels[p][6 * ic + k] = td[6 * isort[ic] + k];
  memcpy(td, c->sub_volt[p], sizeof(double) * 6 * nn);
  for (ic = 0; ic < na; ic++)
    for (k = 0; k < 6; k++) c->sub_volt[p][6 * ic + k] = td[6 * isort[ic] + k];
  memcpy(ti, c->sub_numb[p], sizeof(int64_t) * 6 * nn);
  for (ic = 0; ic < na; ic++)
    for (k = 0; k < 6; k++) c->sub_numb[p][6 * ic + k] = ti[6 * isort[ic] + k];
  memcpy(ti, c->sub_ilevel[p], sizeof(int64_t) * nn);
  for (ic = 0; ic < na; ic++) c->sub_ilevel[p][ic] = ti[isort[ic]];
  c->sub_n[p] = na;
  free(isort); free(ti); free(td);
}

/* checkconductors_work, f3d_mgrid.F:4141-4326: drop points outside the
 * mesh or on conductor interior points, merge duplicate subgrid points. */
static void checkconductors_work(oc_conductors *c, int64_t nx, int64_t ny,
                                 int64_t nz, int64_t mglevel) {
  int64_t nmax = c->sub_nmax[0] > c->sub_nmax[1] ? c->sub_nmax[0] : c->sub_nmax[1];
  int64_t n3 = (nx + 1) * (ny + 1) * (nz + 1);
  int64_t *iii = (int64_t *)calloc((size_t)n3, sizeof(int64_t));
  int64_t ic, p, id;
#define III(ix, iy, iz) iii[(ix) + (nx + 1) * ((iy) + (ny + 1) * (iz))]
  for (ic = 0; ic < c->int_n; ic++) {
    if (c->int_ilevel[ic] != mglevel) continue;
    int64_t ix = c->int_indx[3 * ic], iy = c->int_indx[3 * ic + 1], iz = c->int_indx[3 * ic + 2];
    if (ix < 0 || nx < ix || iy < 0 || ny < iy || iz < 0 || nz < iz) {
      c->int_ilevel[ic] = -1;
      continue;
    }
    if (III(ix, iy, iz) == nmax + 1) {
      c->int_ilevel[ic] = -1;
      continue;
    }
    III(ix, iy, iz) = nmax + 1;
  }
  for (p = 0; p < 2; p++) {
    int64_t *indx = c->sub_indx[p], *ilev = c->sub_ilevel[p], *numb = c->sub_numb[p];
    double *dels = c->sub_dels[p], *volt = c->sub_volt[p];
    for (ic = 0; ic < c->sub_n[p]; ic++) {
      if (ilev[ic] != mglevel) continue;
      int64_t ix = indx[3 * ic], iy = indx[3 * ic + 1], iz = indx[3 * ic + 2];
      if (ix < 0 || nx < ix || iy < 0 || ny < iy || iz < 0 || nz < iz) {
        ilev[ic] = -1;
        continue;
      }
      if (III(ix, iy, iz) == nmax + 1) {
        ilev[ic] = -1;
        continue;
      }
      if (III(ix, iy, iz) == 0) {
        /* stored 1-based, as the Fortran list index */
        III(ix, iy, iz) = ic + 1;
        continue;
      }
      int64_t i = III(ix, iy, iz) - 1;
      if (i < 0) {
        III(ix, iy, iz) = ic + 1;
        continue;
      }
      if (ilev[ic] != ilev[i]) {
        III(ix, iy, iz) = ic + 1;
        continue;
      }
      for (id = 0; id < 6; id++) {
        if (fabs(dels[6 * ic + id]) < 1. && fabs(dels[6 * ic + id]) < fabs(dels[6 * i + id])) {
          dels[6 * i + id] = dels[6 * ic + id];
          volt[6 * i + id] = volt[6 * ic + id];
          numb[6 * i + id] = numb[6 * ic + id];
        }
      }
      ilev[ic] = -1;
    }
  }
#undef III
  free(iii);
}

/* checkconductors, f3d_mgrid.F:3964-4073. */
void oc_checkconductors(int64_t nx, int64_t ny, int64_t nz,
                        double dx, double dy, double dz, oc_conductors *c) {
  oc_getmglevels(nx, ny, nz, dx, dy, dz, c);
  int64_t il;
  for (il = 0; il < c->levels; il++)
    checkconductors_work(c, c->levelnx[il], c->levelny[il], c->levelnz[il], il);
  int64_t l;
  for (l = 0; l <= OC_MAXLEVELS; l++) {
    c->int_istart[l] = 0;
    c->sub_istart[0][l] = 0;
    c->sub_istart[1][l] = 0;
  }
  if (c->int_n > 0) sort_interior(c);
  if (c->sub_n[0] > 0) sort_subgrid(c, 0);
  if (c->sub_n[1] > 0) sort_subgrid(c, 1);
  /* Empty lists: every level range is empty. */
  if (c->int_n == 0) for (l = 0; l <= c->levels; l++) c->int_istart[l] = 0;
  for (int p = 0; p < 2; p++)
    if (c->sub_n[p] == 0) for (l = 0; l <= c->levels; l++) c->sub_istart[p][l] = 0;
}

/* cond_potmg, f3d_mgrid.F:2426-2472. */
static void cond_potmg(oc_conductors *c, int64_t nx, int64_t ny, double *phi,
                       int64_t mglevel, int64_t mgform, int64_t mgform2init) {
  int64_t ic;
  for (ic = c->int_istart[mglevel]; ic < c->int_istart[mglevel + 1]; ic++) {
    int64_t ix = c->int_indx[3 * ic], iy = c->int_indx[3 * ic + 1], iz = c->int_indx[3 * ic + 2];
    if (mglevel == 0 && (mgform == 1 || mgform2init))
      PHI(ix, iy, iz) = c->int_volt[ic];
    else
      PHI(ix, iy, iz) = 0.;
  }
}

/* cond_potmgres, f3d_mgrid.F:2474-2506. */
static void cond_potmgres(oc_conductors *c, int64_t nx, int64_t ny, double *res,
                          int64_t mglevel) {
  int64_t ic;
  for (ic = c->int_istart[mglevel]; ic < c->int_istart[mglevel + 1]; ic++) {
    int64_t ix = c->int_indx[3 * ic], iy = c->int_indx[3 * ic + 1], iz = c->int_indx[3 * ic + 2];
    RES(ix, iy, iz) = 0.;
  }
}

/* handlesubgrid2, f3d_mgrid.F:2628-2764.  d and v point at the six dels
 * and volts of one subgrid point. */
static void handlesubgrid2(int im, int ip, double *pm, double *pp, double dsqicos,
                           const double *d, const double *v, double voltfac,
                           double *denom, double *ppp) {
  double adel;
  if (0. < d[im] && d[im] < +1. && 0. < d[ip] && d[ip] < +1.) {
    *pm = voltfac * v[im] * 2. / (d[im] * (d[ip] + d[im]));
    *pp = voltfac * v[ip] * 2. / (d[ip] * (d[ip] + d[im]));
    *denom = *denom + 2. * (1. - d[ip] * d[im]) / (d[ip] * d[im]) * dsqicos;
    *ppp = dmin(*ppp, d[im]);
    *ppp = dmin(*ppp, d[ip]);
  } else if (0. >= d[im] && d[im] > -1. && 0. >= d[ip] && d[ip] > -1.) {
    *pm = 0.;
    *pp = 0.;
    *denom = *denom - 2. * dsqicos;
    if (fabs(d[im]) > 0.) *ppp = dmin(*ppp, fabs(d[im]));
    else *ppp = dmin(*ppp, 1. - 1.e-9);
    if (fabs(d[ip]) > 0.) *ppp = dmin(*ppp, fabs(d[ip]));
    else *ppp = dmin(*ppp, 1. - 1.e-9);
  } else if (0. < d[im] && d[im] < +1. && 0. >= d[ip] && d[ip] > -1.) {
    adel = fabs(d[ip]);
    *pm = voltfac * v[im] / (d[im] * (adel + 0.5 * d[im]));
    *pp = 0.;
    *denom = *denom + (1. - 2. * d[im] * adel - d[im]) /
                          (d[im] * (adel + 0.5 * d[im])) * dsqicos;
    *ppp = dmin(*ppp, d[im]);
    if (adel > 0.) *ppp = dmin(*ppp, adel);
    else *ppp = dmin(*ppp, 1. - 1.e-9);
  } else if (0. >= d[im] && d[im] > -1. && 0. < d[ip] && d[ip] < +1.) {
    adel = fabs(d[im]);
    *pm = 0.;
    *pp = voltfac * v[ip] / (d[ip] * (adel + 0.5 * d[ip]));
    *denom = *denom + (1. - 2. * d[ip] * adel - d[ip]) /
                          (d[ip] * (adel + 0.5 * d[ip])) * dsqicos;
    if (adel > 0.) *ppp = dmin(*ppp, adel);
    else *ppp = dmin(*ppp, 1. - 1.e-9);
    *ppp = dmin(*ppp, d[ip]);
  } else {
    if (0. < d[im] && d[im] < +1.) {
      *pm = voltfac * v[im] * 2. / (d[im] * (1. + d[im]));
      *pp = *pp * (2. / (1. + d[im]));
      *denom = *denom + 2. * (1. - d[im]) / d[im] * dsqicos;
      *ppp = dmin(*ppp, d[im]);
    } else if (0. >= d[im] && d[im] > -1.) {
      adel = fabs(d[im]);
      *pm = 0.;
      *pp = *pp / (0.5 + adel);
      *denom = *denom - 2. * adel / (adel + 0.5) * dsqicos;
      if (adel > 0.) *ppp = dmin(*ppp, adel);
      else *ppp = dmin(*ppp, 1. - 1.e-9);
    }
    if (0. < d[ip] && d[ip] < +1.) {
      *pp = voltfac * v[ip] * 2. / (d[ip] * (1. + d[ip]));
      *pm = *pm * (2. / (1. + d[ip]));
      *denom = *denom + 2. * (1. - d[ip]) / d[ip] * dsqicos;
      *ppp = dmin(*ppp, d[ip]);
    } else if (0. >= d[ip] && d[ip] > -1.) {
      adel = fabs(d[ip]);
      *pp = 0.;
      *pm = *pm / (0.5 + adel);
      *denom = *denom - 2. * adel / (adel + 0.5) * dsqicos;
      if (adel > 0.) *ppp = dmin(*ppp, adel);
      else *ppp = dmin(*ppp, 1. - 1.e-9);
    }
  }
}

/* condbndymg, f3d_mgrid.F:2766-2980 (icndbndy 1 and 2). */
static void condbndymg(oc_conductors *c, int p, int64_t nx, int64_t ny, int64_t nz,
                       double *phi, const double *rho, double dxsqic, double dysqic,
                       double dzsqic, double spm1, double mgparam,
                       const int64_t *lb, int64_t mglevel, int64_t mgform,
                       int64_t icndbndy) {
  double dxsqicos = dxsqic / mgparam;
  double dysqicos = dysqic / mgparam;
  double dzsqicos = dzsqic / mgparam;
  const int64_t *indx = c->sub_indx[p];
  const double *dels = c->sub_dels[p], *volt = c->sub_volt[p];
  const double *prevphi = c->sub_prevphi[p];
  int64_t ic;
  for (ic = c->sub_istart[p][mglevel]; ic < c->sub_istart[p][mglevel + 1]; ic++) {
    int64_t ix = indx[3 * ic], iy = indx[3 * ic + 1], iz = indx[3 * ic + 2];
    if (ix == 0 && lb[0] < 1) continue;
    if (ix == nx && lb[1] < 1) continue;
    if (iy == 0 && lb[2] < 1) continue;
    if (iy == ny && lb[3] < 1) continue;
    if (iz == 0 && lb[4] < 1) continue;
    if (iz == nz && lb[5] < 1) continue;
    int64_t iym1 = iy - 1, iyp1 = iy + 1;
    if (ny == 0) { iym1 = iy; iyp1 = iy; }
    double pxm = PHI(ix - 1, iy, iz), pxp = PHI(ix + 1, iy, iz);
    double pym = PHI(ix, iym1, iz), pyp = PHI(ix, iyp1, iz);
    double pzm = PHI(ix, iy, iz - 1), pzp = PHI(ix, iy, iz + 1);
    double denom = 1.;
    double voltfac = (mglevel == 0 && mgform == 1) ? 1. : 0.;
    const double *d = dels + 6 * ic, *v = volt + 6 * ic;
    if (icndbndy == 1) {
      int dosubgrid = 0;
      if (0 < d[0] && d[0] < 1.) { pxm = voltfac * v[0] / d[0]; denom = denom + (1. - d[0]) / d[0] * dxsqicos; dosubgrid = 1; }
      if (0 < d[1] && d[1] < 1.) { pxp = voltfac * v[1] / d[1]; denom = denom + (1. - d[1]) / d[1] * dxsqicos; dosubgrid = 1; }
      if (0 < d[2] && d[2] < 1. && ny > 0) { pym = voltfac * v[2] / d[2]; denom = denom + (1. - d[2]) / d[2] * dysqicos; dosubgrid = 1; }
      if (0 < d[3] && d[3] < 1. && ny > 0) { pyp = voltfac * v[3] / d[3]; denom = denom + (1. - d[3]) / d[3] * dysqicos; dosubgrid = 1; }
      if (0 < d[4] && d[4] < 1.) { pzm = voltfac * v[4] / d[4]; denom = denom + (1. - d[4]) / d[4] * dzsqicos; dosubgrid = 1; }
      if (0 < d[5] && d[5] < 1.) { pzp = voltfac * v[5] / d[5]; denom = denom + (1. - d[5]) / d[5] * dzsqicos; dosubgrid = 1; }
      if (dosubgrid)
        PHI(ix, iy, iz) = (RHO(ix, iy, iz) + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic +
                           (pzm + pzp) * dzsqic) / denom + spm1 * prevphi[ic];
    } else if (icndbndy == 2) {
      double ppp = 1.;
      handlesubgrid2(0, 1, &pxm, &pxp, dxsqicos, d, v, voltfac, &denom, &ppp);
      if (ny > 0) handlesubgrid2(2, 3, &pym, &pyp, dysqicos, d, v, voltfac, &denom, &ppp);
      handlesubgrid2(4, 5, &pzm, &pzp, dzsqicos, d, v, voltfac, &denom, &ppp);
      if (ppp < 1.)
        PHI(ix, iy, iz) = (RHO(ix, iy, iz) + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic +
                           (pzm + pzp) * dzsqic) / denom + spm1 * prevphi[ic];
    }
  }
}

/* condbndymgint, f3d_mgrid.F:2982-3092. */
static void condbndymgint(oc_conductors *c, int64_t nx, int64_t ny, int64_t nz,
                          double *phi, const int64_t *lb, int64_t mglevel,
                          int64_t icndbndy) {
  if (icndbndy != 2) return;
  int p;
  for (p = 0; p < 2; p++) {
    const int64_t *indx = c->sub_indx[p];
    const double *dels = c->sub_dels[p];
    int64_t ic;
    for (ic = c->sub_istart[p][mglevel]; ic < c->sub_istart[p][mglevel + 1]; ic++) {
      int64_t ix = indx[3 * ic], iy = indx[3 * ic + 1], iz = indx[3 * ic + 2];
      int64_t ixp1 = ix + 1, ixm1 = ix - 1, iyp1 = iy + 1, iym1 = iy - 1, izp1 = iz + 1, izm1 = iz - 1;
      const double *d = dels + 6 * ic;
      if (!((ixm1 == 0 && lb[0] == 0) || (ixp1 == nx && lb[1] == 0) ||
            (ix == 0 && lb[0] < 1) || (ix == nx && lb[1] < 1))) {
        if (0. >= d[0] && d[0] > -1.) PHI(ixm1, iy, iz) = PHI(ix, iy, iz);
        if (0. >= d[1] && d[1] > -1.) PHI(ixp1, iy, iz) = PHI(ix, iy, iz);
      }
      if (!((iym1 == 0 && lb[2] == 0) || (iyp1 == ny && lb[3] == 0) ||
            (iy == 0 && lb[2] < 1) || (iy == ny && lb[3] < 1))) {
        if (0. >= d[2] && d[2] > -1.) PHI(ix, iym1, iz) = PHI(ix, iy, iz);
        if (0. >= d[3] && d[3] > -1.) PHI(ix, iyp1, iz) = PHI(ix, iy, iz);
      }
      if (!((izm1 == 0 && lb[4] == 0) || (izp1 == nz && lb[5] == 0) ||
            (iz == 0 && lb[4] < 1) || (iz == nz && lb[5] < 1))) {
        if (0. >= d[4] && d[4] > -1.) PHI(ix, iy, izm1) = PHI(ix, iy, iz);
        if (0. >= d[5] && d[5] > -1.) PHI(ix, iy, izp1) = PHI(ix, iy, iz);
      }
    }
  }
}

/* condbndyres, f3d_mgrid.F:3094-3365. */
static void condbndyres(oc_conductors *c, int64_t nx, int64_t ny, int64_t nz,
                        const double *phi, const double *rho, double *res,
                        double dxsqi, double dysqi, double dzsqi, double mgparam,
                        const int64_t *lb, int64_t mglevel, int64_t mgform,
                        int64_t mgform2init, int64_t icndbndy) {
  double cst = 0.5 / (dxsqi + dysqi + dzsqi);
  double dxsqic = dxsqi * cst, dysqic = dysqi * cst, dzsqic = dzsqi * cst;
  double dxsqics = mgparam * dxsqi * cst, dysqics = mgparam * dysqi * cst;
  double dzsqics = mgparam * dzsqi * cst;
  int p;
  for (p = 0; p < 2; p++) {
    const int64_t *indx = c->sub_indx[p];
    const double *dels = c->sub_dels[p], *volt = c->sub_volt[p];
    int64_t ic;
    for (ic = c->sub_istart[p][mglevel]; ic < c->sub_istart[p][mglevel + 1]; ic++) {
      int64_t ix = indx[3 * ic], iy = indx[3 * ic + 1], iz = indx[3 * ic + 2];
      if (ix == 0 && lb[0] == 0) continue;
      if (ix == nx && lb[1] == 0) continue;
      if (iy == 0 && lb[2] == 0) continue;
      if (iy == ny && lb[3] == 0) continue;
      if (iz == 0 && lb[4] == 0) continue;
      if (iz == nz && lb[5] == 0) continue;
      int64_t ixp1 = ix + 1, ixm1 = ix - 1, iyp1 = iy + 1, iym1 = iy - 1, izp1 = iz + 1, izm1 = iz - 1;
      if (ny == 0) { iym1 = iy; iyp1 = iy; }
      double pxm = PHI(ixm1, iy, iz), pxp = PHI(ixp1, iy, iz);
      double pym = PHI(ix, iym1, iz), pyp = PHI(ix, iyp1, iz);
      double pzm = PHI(ix, iy, izm1), pzp = PHI(ix, iy, izp1);
      double denom = 1., ppp = 1.;
      double voltfac = (mglevel == 0 && (mgform == 1 || mgform2init)) ? 1. : 0.;
      const double *d = dels + 6 * ic, *v = volt + 6 * ic;
      if (icndbndy == 1) {
        if (0 < d[0] && d[0] < 1.) { pxm = voltfac * v[0] / d[0]; denom = denom + (1. - d[0]) / d[0] * dxsqic; ppp = dmin(ppp, d[0]); }
        if (0 < d[1] && d[1] < 1.) { pxp = voltfac * v[1] / d[1]; denom = denom + (1. - d[1]) / d[1] * dxsqic; ppp = dmin(ppp, d[1]); }
        if (0 < d[2] && d[2] < 1. && ny > 0) { pym = voltfac * v[2] / d[2]; denom = denom + (1. - d[2]) / d[2] * dysqic; ppp = dmin(ppp, d[2]); }
        if (0 < d[3] && d[3] < 1. && ny > 0) { pyp = voltfac * v[3] / d[3]; denom = denom + (1. - d[3]) / d[3] * dysqic; ppp = dmin(ppp, d[3]); }
        if (0 < d[4] && d[4] < 1.) { pzm = voltfac * v[4] / d[4]; denom = denom + (1. - d[4]) / d[4] * dzsqic; ppp = dmin(ppp, d[4]); }
        if (0 < d[5] && d[5] < 1.) { pzp = voltfac * v[5] / d[5]; denom = denom + (1. - d[5]) / d[5] * dzsqic; ppp = dmin(ppp, d[5]); }
        if (ppp < 1.) {
          if (mgform2init) ppp = 1.;
          RES(ix, iy, iz) = ppp * (RHO(ix, iy, iz) + (pxm + pxp) * dxsqics + (pym + pyp) * dysqics +
                                   (pzm + pzp) * dzsqics - PHI(ix, iy, iz) * mgparam * denom);
        }
      } else if (icndbndy == 2) {
        handlesubgrid2(0, 1, &pxm, &pxp, dxsqic, d, v, voltfac, &denom, &ppp);
        if (ny > 0) handlesubgrid2(2, 3, &pym, &pyp, dysqic, d, v, voltfac, &denom, &ppp);
        handlesubgrid2(4, 5, &pzm, &pzp, dzsqic, d, v, voltfac, &denom, &ppp);
        if (ppp < 1.) {
          if (mgform2init) ppp = 1.;
          RES(ix, iy, iz) = ppp * (RHO(ix, iy, iz) + (pxm + pxp) * dxsqics + (pym + pyp) * dysqics +
                                   (pzm + pzp) * dzsqics - PHI(ix, iy, iz) * mgparam * denom);
          if (0. >= d[0] && d[0] >= -1.) RES(ixm1, iy, iz) = 0.;
          if (0. >= d[1] && d[1] >= -1.) RES(ixp1, iy, iz) = 0.;
          if (0. >= d[2] && d[2] >= -1. && ny > 0) RES(ix, iym1, iz) = 0.;
          if (0. >= d[3] && d[3] >= -1. && ny > 0) RES(ix, iyp1, iz) = 0.;
          if (0. >= d[4] && d[4] >= -1.) RES(ix, iy, izm1) = 0.;
          if (0. >= d[5] && d[5] >= -1.) RES(ix, iy, izp1) = 0.;
        }
      }
    }
  }
}

/* sorhalfpass3d, f3d_mgrid.F:2095-2291 (lprecalccoeffs=.false., no bends). */
void oc_sorhalfpass3d(int64_t parity, int64_t mglevel,
                      int64_t nx, int64_t ny, int64_t nz,
                      double *phi, const double *rho,
                      double dxsqi, double dysqi, double dzsqi,
                      const int64_t *lb, double mgparam, int64_t mgform,
                      int64_t lcndbndy, int64_t icndbndy, oc_conductors *c) {
  double cst = mgparam * 0.5 / (dxsqi + dysqi + dzsqi);
  double dxsqic = dxsqi * cst, dysqic = dysqi * cst, dzsqic = dzsqi * cst;
  double spm1 = 1. - mgparam;
  int64_t ixmin = 0, ixmax = nx, iymin = 0, iymax = ny, izmin = 0, izmax = nz;
  if (lb[0] <= 0) ixmin = 1;
  if (lb[1] <= 0) ixmax = nx - 1;
  if (lb[2] <= 0) iymin = 1;
  if (lb[3] <= 0) iymax = ny - 1;
  if (lb[4] <= 0) izmin = 1;
  if (lb[5] <= 0) izmax = nz - 1;
  int p = (int)parity;
  int64_t ic, ix, iy, iz;
  if (lcndbndy && c) {
    for (ic = c->sub_istart[p][mglevel]; ic < c->sub_istart[p][mglevel + 1]; ic++) {
      const int64_t *q = c->sub_indx[p] + 3 * ic;
      c->sub_prevphi[p][ic] = PHI(q[0], q[1], q[2]);
    }
  }
  OC_PARFOR(private(iy, ix))
  for (iz = izmin; iz <= izmax; iz++)
    for (iy = iymin; iy <= iymax; iy++) {
      int64_t ix1 = ixmin + (ixmin + iy + iz + parity) % 2;
      for (ix = ix1; ix <= ixmax; ix += 2)
        PHI(ix, iy, iz) = RHO(ix, iy, iz) +
                          (PHI(ix - 1, iy, iz) + PHI(ix + 1, iy, iz)) * dxsqic +
                          (PHI(ix, iy - 1, iz) + PHI(ix, iy + 1, iz)) * dysqic +
                          (PHI(ix, iy, iz - 1) + PHI(ix, iy, iz + 1)) * dzsqic +
                          spm1 * PHI(ix, iy, iz);
    }
  if (c) {
    if (lcndbndy)
      condbndymg(c, p, nx, ny, nz, phi, rho, dxsqic, dysqic, dzsqic, spm1, mgparam,
                 lb, mglevel, mgform, icndbndy);
    cond_potmg(c, nx, ny, phi, mglevel, mgform, 0);
    condbndymgint(c, nx, ny, nz, phi, lb, mglevel, icndbndy);
  }
}

/* sorpass3d, f3d_mgrid.F:1949-2093 (serial: parity 0 then 1). */
void oc_sorpass3d(int64_t mglevel, int64_t nx, int64_t ny, int64_t nz,
                  double *phi, const double *rho,
                  double dxsqi, double dysqi, double dzsqi,
                  const int64_t *lb, double mgparam, int64_t mgform,
                  int64_t lcndbndy, int64_t icndbndy, oc_conductors *c) {
  if (c) {
    cond_potmg(c, nx, ny, phi, mglevel, mgform, 0);
    condbndymgint(c, nx, ny, nz, phi, lb, mglevel, icndbndy);
  }
  int64_t parity;
  for (parity = 0; parity <= 1; parity++) {
    oc_sorhalfpass3d(parity, mglevel, nx, ny, nz, phi, rho, dxsqi, dysqi, dzsqi, lb,
                     mgparam, mgform, lcndbndy, icndbndy, c);
    oc_applyboundaryconditions3d(nx, ny, nz, 1, 1, 1, phi, lb, 0, 0);
  }
}

/* residual3d, f3d_mgrid.F:3367-3598 (lprecalccoeffs=.false., guard 1). */
void oc_residual3d(int64_t nx, int64_t ny, int64_t nz,
                   double dxsqi, double dysqi, double dzsqi,
                   const double *phi, const double *rho, double *res,
                   int64_t mglevel, const int64_t *lb, double mgparam,
                   int64_t mgform, int64_t mgform2init,
                   int64_t lcndbndy, int64_t icndbndy, oc_conductors *c) {
  const int64_t g = 1;
  double cst = 0.5 / (dxsqi + dysqi + dzsqi);
  double dxsqic = dxsqi * mgparam * cst, dysqic = dysqi * mgparam * cst;
  double dzsqic = dzsqi * mgparam * cst;
  int64_t ixmin = 0, ixmax = nx, iymin = 0, iymax = ny, izmin = 0, izmax = nz;
  if (lb[0] == 0) ixmin = 1;
  if (lb[1] == 0) ixmax = nx - 1;
  if (lb[2] == 0) iymin = 1;
  if (lb[3] == 0) iymax = ny - 1;
  if (lb[4] == 0) izmin = 1;
  if (lb[5] == 0) izmax = nz - 1;
  int64_t ix, iy, iz;
  if (lb[4] == 0)
    for (iz = -g; iz <= 0; iz++)
      for (iy = -g; iy <= ny + g; iy++)
        for (ix = -g; ix <= nx + g; ix++) RES(ix, iy, iz) = 0.;
  OC_PARFOR(private(iy, ix))
  for (iz = izmin; iz <= izmax; iz++) {
    if (lb[2] == 0)
      for (iy = -g; iy <= 0; iy++)
        for (ix = -g; ix <= nx + g; ix++) RES(ix, iy, iz) = 0.;
    for (iy = iymin; iy <= iymax; iy++) {
      if (lb[0] == 0)
        for (ix = -g; ix <= 0; ix++) RES(ix, iy, iz) = 0.;
      for (ix = ixmin; ix <= ixmax; ix++)
        RES(ix, iy, iz) = RHO(ix, iy, iz) +
                          (PHI(ix - 1, iy, iz) + PHI(ix + 1, iy, iz)) * dxsqic +
                          (PHI(ix, iy - 1, iz) + PHI(ix, iy + 1, iz)) * dysqic +
                          (PHI(ix, iy, iz - 1) + PHI(ix, iy, iz + 1)) * dzsqic -
                          PHI(ix, iy, iz) * mgparam;
      if (lb[1] == 0)
        for (ix = nx; ix <= nx + g; ix++) RES(ix, iy, iz) = 0.;
    }
    if (lb[3] == 0)
      for (iy = ny; iy <= ny + g; iy++)
        for (ix = -g; ix <= nx + g; ix++) RES(ix, iy, iz) = 0.;
  }
  if (lb[5] == 0)
    for (iz = nz; iz <= nz + g; iz++)
      for (iy = -g; iy <= ny + g; iy++)
        for (ix = -g; ix <= nx + g; ix++) RES(ix, iy, iz) = 0.;
  if (c) {
    cond_potmgres(c, nx, ny, res, mglevel);
    if (lcndbndy)
      condbndyres(c, nx, ny, nz, phi, rho, res, dxsqi, dysqi, dzsqi, mgparam, lb,
                  mglevel, mgform, mgform2init, icndbndy);
  }
  oc_applyboundaryconditions3d(nx, ny, nz, g, g, g, res, lb, 0, 0);
}

/* restrict3d, f3d_mgrid.F:1593-1775 (serial, offsets 0, res guard 1). */
void oc_restrict3d(int64_t nx, int64_t ny, int64_t nz, const double *res,
                   int64_t nxc, int64_t nyc, int64_t nzc, double *rhocoarse,
                   double ff, const int64_t *lbc) {
  const int64_t g = 1, nw = 2 * MGCOARSENING;
  int64_t ixcmin = 0, ixcmax = nxc, iycmin = 0, iycmax = nyc, izcmin = 0, izcmax = nzc;
  if (lbc[0] == 0) ixcmin = 1;
  if (lbc[1] == 0) ixcmax = nxc - 1;
  if (lbc[2] == 0) iycmin = 1;
  if (lbc[3] == 0) iycmax = nyc - 1;
  if (lbc[4] == 0) izcmin = 1;
  if (lbc[5] == 0) izcmax = nzc - 1;
  double dxi = 1. * nxc / nx;
  double dyi = ny > 0 ? 1. * nyc / ny : 1.;
  double dzi = 1. * nzc / nz;
  int64_t *ixmina = (int64_t *)calloc(nxc + 1, sizeof(int64_t));
  int64_t *ixmaxa = (int64_t *)calloc(nxc + 1, sizeof(int64_t));
  int64_t *iymina = (int64_t *)calloc(nyc + 1, sizeof(int64_t));
  int64_t *iymaxa = (int64_t *)calloc(nyc + 1, sizeof(int64_t));
  int64_t *izmina = (int64_t *)calloc(nzc + 1, sizeof(int64_t));
  int64_t *izmaxa = (int64_t *)calloc(nzc + 1, sizeof(int64_t));
  double *wxa = (double *)calloc((nw + 1) * (nxc + 1), sizeof(double));
  double *wya = (double *)calloc((nw + 1) * (nyc + 1), sizeof(double));
  double *wza = (double *)calloc((nw + 1) * (nzc + 1), sizeof(double));
  int64_t ixc, iyc, izc, ix, iy, iz;
  /* lxoffset = 0: (iz + 1.*lzoffset/nzcoarse) is (iz + 0.) */
  for (izc = izcmin; izc <= izcmax; izc++) {
    int64_t izmin = ((izc - 1) * nz + 4 * nzc) / nzc - 3;
    int64_t izmax = ((izc + 1) * nz - 1) / nzc;
    izmina[izc] = izmin > -g ? izmin : -g;
    izmaxa[izc] = izmax < nz + g ? izmax : nz + g;
    for (iz = izmin; iz <= izmax; iz++)
      wza[(iz - izmin) + (nw + 1) * izc] = 1. - fabs(izc - (iz + 0.) * dzi);
  }
  for (iyc = iycmin; iyc <= iycmax; iyc++) {
    int64_t iymin = ((iyc - 1) * ny + 4 * nyc) / nyc - 3;
    int64_t iymax = ((iyc + 1) * ny - 1) / nyc;
    iymina[iyc] = iymin > -g ? iymin : -g;
    iymaxa[iyc] = iymax < ny + g ? iymax : ny + g;
    for (iy = iymin; iy <= iymax; iy++)
      wya[(iy - iymin) + (nw + 1) * iyc] = 1. - fabs(iyc - (iy + 0.) * dyi);
  }
  for (ixc = ixcmin; ixc <= ixcmax; ixc++) {
    int64_t ixmin = ((ixc - 1) * nx + 4 * nxc) / nxc - 3;
    int64_t ixmax = ((ixc + 1) * nx - 1) / nxc;
    ixmina[ixc] = ixmin > -g ? ixmin : -g;
    ixmaxa[ixc] = ixmax < nx + g ? ixmax : nx + g;
    for (ix = ixmin; ix <= ixmax; ix++)
      wxa[(ix - ixmin) + (nw + 1) * ixc] = 1. - fabs(ixc - (ix + 0.) * dxi);
  }
#define RC(i, j, k) rhocoarse[(i) + (nxc + 1) * ((j) + (nyc + 1) * (k))]
  if (lbc[4] == 0)
    for (iyc = 0; iyc <= nyc; iyc++)
      for (ixc = 0; ixc <= nxc; ixc++) RC(ixc, iyc, 0) = 0.;
  OC_PARFOR(private(iyc, ixc, iz, iy, ix))
  for (izc = izcmin; izc <= izcmax; izc++) {
    int64_t izmin = izmina[izc], izmax = izmaxa[izc];
    const double *wz = wza + (nw + 1) * izc;
    if (lbc[2] == 0)
      for (ixc = 0; ixc <= nxc; ixc++) RC(ixc, 0, izc) = 0.;
    for (iyc = iycmin; iyc <= iycmax; iyc++) {
      int64_t iymin = iymina[iyc], iymax = iymaxa[iyc];
      const double *wy = wya + (nw + 1) * iyc;
      if (lbc[0] == 0) RC(0, iyc, izc) = 0.;
      for (ixc = ixcmin; ixc <= ixcmax; ixc++) {
        int64_t ixmin = ixmina[ixc], ixmax = ixmaxa[ixc];
        const double *wx = wxa + (nw + 1) * ixc;
        double r = 0., w = 0.;
        for (iz = izmin; iz <= izmax; iz++)
          for (iy = iymin; iy <= iymax; iy++)
            for (ix = ixmin; ix <= ixmax; ix++) {
              r = r + wx[ix - ixmin] * wy[iy - iymin] * wz[iz - izmin] * RES(ix, iy, iz);
              w = w + wx[ix - ixmin] * wy[iy - iymin] * wz[iz - izmin];
            }
        if (w > 0.) {
          r = r / w;
          RC(ixc, iyc, izc) = r * ff;
        } else {
          RC(ixc, iyc, izc) = 0.;
        }
      }
      if (lbc[1] == 0) RC(nxc, iyc, izc) = 0.;
    }
    if (lbc[3] == 0)
      for (ixc = 0; ixc <= nxc; ixc++) RC(ixc, nyc, izc) = 0.;
  }
  if (lbc[5] == 0)
    for (iyc = 0; iyc <= nyc; iyc++)
      for (ixc = 0; ixc <= nxc; ixc++) RC(ixc, iyc, nzc) = 0.;
#undef RC
  free(ixmina); free(ixmaxa); free(iymina); free(iymaxa); free(izmina); free(izmaxa);
  free(wxa); free(wya); free(wza);
}

/* expand3d, f3d_mgrid.F:1777-1947 (ny > 0, offsets 0, no precalc). */
void oc_expand3d(int64_t nx, int64_t ny, int64_t nz, double *phi,
                 int64_t nxc, int64_t nyc, int64_t nzc, const double *phicoarse,
                 const int64_t *lb) {
  int64_t ixmin = 0, ixmax = nx, iymin = 0, iymax = ny, izmin = 0, izmax = nz;
  if (lb[0] == 0) ixmin = 1;
  if (lb[1] == 0) ixmax = nx - 1;
  if (lb[2] == 0) iymin = 1;
  if (lb[3] == 0) iymax = ny - 1;
  if (lb[4] == 0) izmin = 1;
  if (lb[5] == 0) izmax = nz - 1;
#define PC(i, j, k) phicoarse[IDX(i, j, k, nxc, nyc, 1)]
  int64_t ix, iy, iz;
  OC_PARFOR(private(iy, ix))
  for (iz = izmin; iz <= izmax; iz++) {
    int64_t jz = (iz * nzc) / nz;
    double wz = 1. * (iz * nzc) / nz - jz;
    for (iy = iymin; iy <= iymax; iy++) {
      int64_t jy = (iy * nyc) / ny;
      double wy = 1. * (iy * nyc) / ny - jy;
      for (ix = ixmin; ix <= ixmax; ix++) {
        int64_t jx = (ix * nxc) / nx;
        double wx = 1. * (ix * nxc) / nx - jx;
        PHI(ix, iy, iz) = PHI(ix, iy, iz) +
            (1. - wx) * (1. - wy) * (1. - wz) * PC(jx, jy, jz) +
            wx * (1. - wy) * (1. - wz) * PC(jx + 1, jy, jz) +
            (1. - wx) * wy * (1. - wz) * PC(jx, jy + 1, jz) +
            wx * wy * (1. - wz) * PC(jx + 1, jy + 1, jz) +
            (1. - wx) * (1. - wy) * wz * PC(jx, jy, jz + 1) +
            wx * (1. - wy) * wz * PC(jx + 1, jy, jz + 1) +
            (1. - wx) * wy * wz * PC(jx, jy + 1, jz + 1) +
            wx * wy * wz * PC(jx + 1, jy + 1, jz + 1);
      }
    }
  }
#undef PC
}

/* vcycle, f3d_mgrid.F:1101-1591 (serial, mgscale unused). */
static void vcycle(int64_t mglevel, int64_t nx, int64_t ny, int64_t nz,
                   double dx, double dy, double dz, double *phi, double *rho,
                   const int64_t *bounds, double mgparam, int64_t mgform,
                   int64_t mgmaxlevels, int64_t downpasses, int64_t uppasses,
                   int64_t lcndbndy, int64_t icndbndy, oc_conductors *c) {
  double dxsqi = 1. / (dx * dx), dysqi = 1. / (dy * dy), dzsqi = 1. / (dz * dz);
  int64_t i;
  for (i = 0; i < downpasses; i++)
    oc_sorpass3d(mglevel, nx, ny, nz, phi, rho, dxsqi, dysqi, dzsqi, bounds, mgparam,
                 mgform, lcndbndy, icndbndy, c);
  if (nx >= 4 && ny >= 4 && nz >= 4 && mglevel < mgmaxlevels) {
    int64_t nres = (nx + 3) * (ny + 3) * (nz + 3);
    double *res = (double *)malloc(sizeof(double) * nres);
    /* Fortran leaves res uninitialised; every element read later is
     * written by residual3d, so the fill value is immaterial. */
    OC_PARFOR()
    for (int64_t k = 0; k < nres; k++) res[k] = 0.;
    oc_residual3d(nx, ny, nz, dxsqi, dysqi, dzsqi, phi, rho, res, mglevel, bounds,
                  mgparam, mgform, 0, lcndbndy, icndbndy, c);
    int64_t nxc, nyc, nzc;
    double dxc, dyc, dzc;
    oc_getnextcoarselevel3d(nx, ny, nz, dx, dy, dz, &nxc, &nyc, &nzc, &dxc, &dyc, &dzc);
    double dxcsqi = 1. / (dxc * dxc), dycsqi = 1. / (dyc * dyc), dzcsqi = 1. / (dzc * dzc);
    int64_t npc = (nxc + 3) * (nyc + 3) * (nzc + 3);
    double *phicoarse = (double *)calloc(npc, sizeof(double));
    double *rhocoarse = (double *)calloc((nxc + 1) * (nyc + 1) * (nzc + 1), sizeof(double));
    double ff = (dxsqi + dysqi + dzsqi) / (dxcsqi + dycsqi + dzcsqi);
    oc_restrict3d(nx, ny, nz, res, nxc, nyc, nzc, rhocoarse, ff, bounds);
    vcycle(mglevel + 1, nxc, nyc, nzc, dxc, dyc, dzc, phicoarse, rhocoarse, bounds,
           mgparam, mgform, mgmaxlevels, downpasses, uppasses, lcndbndy, icndbndy, c);
    oc_expand3d(nx, ny, nz, phi, nxc, nyc, nzc, phicoarse, bounds);
    oc_applyboundaryconditions3d(nx, ny, nz, 1, 1, 1, phi, bounds, 0, 0);
    free(phicoarse); free(rhocoarse); free(res);
  }
  for (i = 0; i < uppasses; i++)
    oc_sorpass3d(mglevel, nx, ny, nz, phi, rho, dxsqi, dysqi, dzsqi, bounds, mgparam,
                 mgform, lcndbndy, icndbndy, c);
}

/* multigrid3dsolve, f3d_mgrid.F:83-500 (serial, iwhich=0, no bends,
 * lprecalccoeffs=.false., gridmode=1).  Returns 0, or -1 when the conductor
 * levels do not match the grid. */
int64_t oc_multigrid3dsolve(int64_t nx, int64_t ny, int64_t nz,
                            double dx, double dy, double dz,
                            double *phi, double *rho, const int64_t *bounds,
                            double mgparam, int64_t mgform,
                            int64_t *mgiters, int64_t mgmaxiters,
                            int64_t mgmaxlevels, double *mgerror, double mgtol,
                            int64_t downpasses, int64_t uppasses,
                            int64_t lcndbndy, int64_t icndbndy,
                            oc_conductors *c) {
  double eps0 = oc_eps0();
  double dxsqi = 1. / (dx * dx), dysqi = 1. / (dy * dy), dzsqi = 1. / (dz * dz);
  double reps0c = mgparam / (eps0 * 2. * (dxsqi + dysqi + dzsqi));
  oc_checkconductors(nx, ny, nz, dx, dy, dz, c);
  int64_t nrho = (nx + 1) * (ny + 1) * (nz + 1);
  int64_t nphi = (nx + 3) * (ny + 3) * (nz + 3);
  int64_t k, ix, iy, iz;
  OC_PARFOR()
  for (k = 0; k < nrho; k++) rho[k] = rho[k] * reps0c;
  oc_applyboundaryconditions3d(nx, ny, nz, 1, 1, 1, phi, bounds, 0, 0);
  double *rhosave = NULL, *res = NULL;
  if (mgform == 2) {
    rhosave = (double *)malloc(sizeof(double) * nrho);
    res = (double *)calloc(nphi, sizeof(double));
    memcpy(rhosave, rho, sizeof(double) * nrho);
  }
  double *phisave = (double *)malloc(sizeof(double) * nphi);
  double *planeerr = (double *)malloc(sizeof(double) * (nz + 1));
  int64_t *planenan = (int64_t *)malloc(sizeof(int64_t) * (nz + 1));
  *mgiters = 0;
  *mgerror = 2. * mgtol + 1.;
  while (*mgerror > mgtol && *mgiters < mgmaxiters) {
    *mgiters = *mgiters + 1;
    OC_PARFOR()
    for (k = 0; k < nphi; k++) phisave[k] = phi[k];
    if (mgform == 2) {
      cond_potmg(c, nx, ny, phisave, 0, mgform, 1);
      condbndymgint(c, nx, ny, nz, phisave, bounds, 0, icndbndy);
      oc_residual3d(nx, ny, nz, dxsqi, dysqi, dzsqi, phisave, rhosave, res, 0, bounds,
                    mgparam, mgform, 1, lcndbndy, icndbndy, c);
      for (iz = 0; iz <= nz; iz++)
        for (iy = 0; iy <= ny; iy++)
          for (ix = 0; ix <= nx; ix++) RHO(ix, iy, iz) = RES(ix, iy, iz);
      for (k = 0; k < nphi; k++) phi[k] = 0.;
    }
    vcycle(0, nx, ny, nz, dx, dy, dz, phi, rho, bounds, mgparam, mgform, mgmaxlevels,
           downpasses, uppasses, lcndbndy, icndbndy, c);
    if (mgform == 2) {
      for (k = 0; k < nphi; k++) phi[k] = phi[k] + phisave[k];
      oc_applyboundaryconditions3d(nx, ny, nz, 1, 1, 1, phi, bounds, 0, 0);
    }
    /* err = err > d ? err : d over the nodes in order.  Each plane is
     * folded from 0 in parallel; a NaN restarts the fold, so a plane that
     * holds one gives its own fold, any other plane its maximum with the
     * running value: the serial loop's result, NaNs included. */
    double err = 0.;
    OC_PARFOR(private(iy, ix))
    for (iz = 0; iz <= nz; iz++) {
      double pe = 0.;
      int64_t nan = 0;
      for (iy = 0; iy <= ny; iy++)
        for (ix = 0; ix <= nx; ix++) {
          double d = fabs(phisave[IDX(ix, iy, iz, nx, ny, 1)] - PHI(ix, iy, iz));
          if (d != d) nan = 1;
          pe = pe > d ? pe : d;
        }
      planeerr[iz] = pe;
      planenan[iz] = nan;
    }
    for (iz = 0; iz <= nz; iz++)
      err = planenan[iz] ? planeerr[iz] : (err > planeerr[iz] ? err : planeerr[iz]);
    *mgerror = err;
  }
  oc_applyboundaryconditions3d(nx, ny, nz, 1, 1, 1, phi, bounds, 1, 0);
  if (mgform == 2) {
    memcpy(rho, rhosave, sizeof(double) * nrho);
    free(rhosave);
    free(res);
  }
  free(phisave);
  free(planeerr);
  free(planenan);
  reps0c = 1. / reps0c;
  OC_PARFOR()
  for (k = 0; k < nrho; k++) rho[k] = rho[k] * reps0c;
  return 0;
}
