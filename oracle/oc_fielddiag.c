/*
 * oc_fielddiag.c -- TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline).
 *
 * FP64 restatement of the two field diagnostics MultiGrid3D computes on
 * special steps (scripts/multigrid.py:433-467): the line charge
 * gtlchg3dfromrho (pywarp90/dw3d.F:229-320) and the electrostatic energy
 * getese3dfromrhophi (dw3d.F:708-768), serial case (one processor, no MPI
 * reduction).  rho has no guard cells, phi one (nxguardphi = 1).  Every sum
 * runs in the order the Fortran loops and the SUM intrinsic give.  Compile
 * with -ffp-contract=off.
 */
#include "warp_oracle.h"

#define IDXR(ix, iy, iz) \
  ((int64_t)(ix) + (nx + 1) * ((int64_t)(iy) + (ny + 1) * (int64_t)(iz)))
#define IDXP(ix, iy, iz) \
  ((int64_t)((ix) + 1) + (nx + 3) * ((int64_t)((iy) + 1) + (ny + 3) * (int64_t)((iz) + 1)))

/* dw3d.F:254-279: charge summed over each z plane; with l2symtry only
 * iy >= 0 is stored and rows iy >= 1 count twice, with l4symtry only the
 * quadrant ix, iy >= 0 is stored. */
void oc_plane_sums(const double *rho, int64_t nx, int64_t ny, int64_t nz,
                   int64_t l2symtry, int64_t l4symtry, double *scrtch) {
  int64_t iz;
  OC_PARFOR()
  for (iz = 0; iz <= nz; iz++) {
    int64_t ix, iy;
    if (l2symtry) {
      double a = 0., b = 0.;
      for (ix = 0; ix <= nx; ix++) a = a + rho[IDXR(ix, 0, iz)];
      for (iy = 1; iy <= ny; iy++)
        for (ix = 0; ix <= nx; ix++) b = b + rho[IDXR(ix, iy, iz)];
      scrtch[iz] = a + 2. * b;
    } else if (l4symtry) {
      double s1 = 0., s2 = 0., s3 = 0.;
      for (iy = 1; iy <= ny; iy++)
        for (ix = 1; ix <= nx; ix++) s1 = s1 + rho[IDXR(ix, iy, iz)];
      for (ix = 1; ix <= nx; ix++) s2 = s2 + rho[IDXR(ix, 0, iz)];
      for (iy = 1; iy <= ny; iy++) s3 = s3 + rho[IDXR(0, iy, iz)];
      scrtch[iz] = rho[IDXR(0, 0, iz)] + 4. * s1 + 2. * s2 + 2. * s3;
    } else {
      double a = 0.;
      for (iy = 0; iy <= ny; iy++)
        for (ix = 0; ix <= nx; ix++) a = a + rho[IDXR(ix, iy, iz)];
      scrtch[iz] = a;
    }
  }
}

/* dw3d.F:282-316: the plane sums times dx*dy, linearly interpolated onto
 * the nzzarr+1 points of the Z_arrays mesh (zzmin + iz*dzz, in the beam
 * frame).  scrtch holds nz+1 plane sums. */
void oc_linechg_from_planes(const double *scrtch, int64_t nz, double dx, double dy,
                            double dz, double zgrid, double zmmin, double zzmin,
                            double zbeam, double dzz, int64_t nzzarr, double *linechg) {
  double dzi = (nz == 0) ? 0. : 1. / dz;
  double zz = zgrid + zmmin - zzmin - zbeam;
  int64_t iz;
  for (iz = 0; iz <= nzzarr; iz++) {
    double zg = (iz * dzz - zz) * dzi;
    int64_t izg = (int64_t)(zg + nz) - nz;
    double wzg = zg - izg;
    if (0 <= izg && izg <= nz)
      linechg[iz] = scrtch[izg] * dx * dy * (1. - wzg);
    else
      linechg[iz] = 0.;
    if (0 < izg + 1 && izg + 1 <= nz)
      linechg[iz] = linechg[iz] + scrtch[izg + 1] * dx * dy * wzg;
  }
}

/* gtlchg3dfromrho, dw3d.F:229-320, for the last (only) processor. */
void oc_gtlchg3d(const double *rho, int64_t nx, int64_t ny, int64_t nz, double dx,
                 double dy, double dz, double zgrid, double zmmin, double zzmin,
                 double zbeam, double dzz, int64_t nzzarr, int64_t l2symtry,
                 int64_t l4symtry, double *scrtch, double *linechg) {
  oc_plane_sums(rho, nx, ny, nz, l2symtry, l4symtry, scrtch);
  oc_linechg_from_planes(scrtch, nz, dx, dy, dz, zgrid, zmmin, zzmin, zbeam, dzz, nzzarr,
                         linechg);
}

/* getese3dfromrhophi, dw3d.F:708-768: rho*phi summed over z for each
 * (ix, iy) column, the symmetry factors, the columns summed over iy, then
 * over ix, times dx*dy*dz/2.  Returns ese (the reference adds into an ese
 * that its caller zeroes, multigrid.py:454). */
double oc_getese3d(const double *rho, const double *phi, int64_t nx, int64_t ny, int64_t nz,
                   double dx, double dy, double dz, int64_t l4symtry, int64_t l2symtry) {
  double *s = (double *)calloc((size_t)((nx + 1) * (ny + 1)), sizeof(double));
  double ese = 0.;
  int64_t ix, iy, iz;
  if (!s) return NAN;
  OC_PARFOR(private(ix, iz))
  for (iy = 0; iy <= ny; iy++)
    for (iz = 0; iz <= nz; iz++)
      for (ix = 0; ix <= nx; ix++)
        s[ix + (nx + 1) * iy] = s[ix + (nx + 1) * iy] + rho[IDXR(ix, iy, iz)] * phi[IDXP(ix, iy, iz)];
  if (l4symtry) {
    for (iy = 1; iy <= ny; iy++) s[(nx + 1) * iy] = 2. * s[(nx + 1) * iy];
    for (ix = 1; ix <= nx; ix++) s[ix] = 2. * s[ix];
    for (iy = 1; iy <= ny; iy++)
      for (ix = 1; ix <= nx; ix++) s[ix + (nx + 1) * iy] = 4. * s[ix + (nx + 1) * iy];
  } else if (l2symtry) {
    for (iy = 0; iy <= ny; iy++)
      for (ix = 1; ix <= nx; ix++) s[ix + (nx + 1) * iy] = 2. * s[ix + (nx + 1) * iy];
  }
  for (iy = 1; iy <= ny; iy++)
    for (ix = 0; ix <= nx; ix++) s[ix] = s[ix] + s[ix + (nx + 1) * iy];
  for (ix = 0; ix <= nx; ix++) ese = ese + s[ix];
  ese = dx * dy * dz * 0.5 * ese;
  free(s);
  return ese;
}
