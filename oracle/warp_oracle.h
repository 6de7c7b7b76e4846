/*
 * warp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of WARP's electrostatic PIC hot path, used as the
 * parity oracle by tests/, by __graft_entry__.smoke() and as the timed CPU
 * baseline in bench.py.  Nothing in the product (warp_amd/) may link, load
 * or call this code.
 *
 * Every function follows the reference loop for loop and in the same
 * floating-point order; the citations are to pywarp90/ files of the
 * reference (radiasoft/warp).  Arrays are Fortran-ordered (x fastest) with
 * the reference's guard cells: phi is (-1:nx+1)^3, rho is (0:nx)^3.
 */
#ifndef WARP_ORACLE_H
#define WARP_ORACLE_H
#include <stdint.h>
#include <stdlib.h>
#include <math.h>

/* Loops whose iterations write disjoint elements and read nothing another
 * iteration writes run in parallel under OpenMP; the results are the same
 * bits as the serial loop. */
#define OC_PRAGMA(x) _Pragma(#x)
#ifdef _OPENMP
#define OC_PARFOR(...) OC_PRAGMA(omp parallel for schedule(static) __VA_ARGS__)
#else
#define OC_PARFOR(...)
#endif
void oc_set_threads(int64_t n);
int64_t oc_get_threads(void);

#define OC_MAXLEVELS 101

/* Conductor data, the same model as ConductorType / ConductorInteriorType /
 * ConductorSubGridType (pywarp90/f3d.v:105-164).  Index 0 of the subgrid
 * arrays is the even list, 1 the odd list.  istart[] is 0-based here:
 * points of level l are [istart[l], istart[l+1]). */
typedef struct {
  int64_t levels;
  int64_t levelnx[OC_MAXLEVELS], levelny[OC_MAXLEVELS], levelnz[OC_MAXLEVELS];
  double levellx[OC_MAXLEVELS], levelly[OC_MAXLEVELS], levellz[OC_MAXLEVELS];
  int64_t int_n, int_nmax;
  int64_t *int_indx;            /* 3*nmax */
  double *int_volt;             /* nmax */
  int64_t *int_numb;            /* nmax */
  int64_t *int_ilevel;          /* nmax */
  int64_t int_istart[OC_MAXLEVELS + 1];
  int64_t sub_n[2], sub_nmax[2];
  int64_t *sub_indx[2];         /* 3*nmax */
  double *sub_dels[2];          /* 6*nmax, order mx,px,my,py,mz,pz */
  double *sub_volt[2];          /* 6*nmax */
  int64_t *sub_numb[2];         /* 6*nmax */
  int64_t *sub_ilevel[2];       /* nmax */
  double *sub_prevphi[2];       /* nmax */
  int64_t sub_istart[2][OC_MAXLEVELS + 1];
} oc_conductors;

/* Grid-line intercepts of one object on one MG level, the model of
 * ConductorInterceptType (f3d.v).  x intercepts are (nxicpt, 0:ny, 0:nz),
 * y intercepts (nyicpt, 0:nx, 0:nz), z intercepts (nzicpt, 0:nx, 0:ny),
 * each with the intercept index fastest. */
typedef struct {
  double xmmin, ymmin, zmmin, dx, dy, dz;
  int64_t nx, ny, nz, ix, iy, iz, mglevel;
  int64_t nxicpt, nyicpt, nzicpt;
  double *xintercepts, *yintercepts, *zintercepts;
  double *xvoltages, *yvoltages, *zvoltages;
  int64_t *xcondids, *ycondids, *zcondids;
} oc_intercepts;

#ifdef __cplusplus
extern "C" {
#endif

double oc_eps0(void);

void oc_applyboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz,
                                  int64_t gx, int64_t gy, int64_t gz,
                                  double *u, const int64_t *bounds,
                                  int64_t lwithdirichlet, int64_t lzerodirichlet);

void oc_getnextcoarselevel3d(int64_t nx, int64_t ny, int64_t nz,
                             double dx, double dy, double dz,
                             int64_t *nxc, int64_t *nyc, int64_t *nzc,
                             double *dxc, double *dyc, double *dzc);

void oc_getmglevels(int64_t nx, int64_t ny, int64_t nz,
                    double dx, double dy, double dz, oc_conductors *c);

void oc_checkconductors(int64_t nx, int64_t ny, int64_t nz,
                        double dx, double dy, double dz, oc_conductors *c);

void oc_sorhalfpass3d(int64_t parity, int64_t mglevel,
                      int64_t nx, int64_t ny, int64_t nz,
                      double *phi, const double *rho,
                      double dxsqi, double dysqi, double dzsqi,
                      const int64_t *bounds, double mgparam, int64_t mgform,
                      int64_t lcndbndy, int64_t icndbndy, oc_conductors *c);

void oc_sorpass3d(int64_t mglevel, int64_t nx, int64_t ny, int64_t nz,
                  double *phi, const double *rho,
                  double dxsqi, double dysqi, double dzsqi,
                  const int64_t *bounds, double mgparam, int64_t mgform,
                  int64_t lcndbndy, int64_t icndbndy, oc_conductors *c);

void oc_residual3d(int64_t nx, int64_t ny, int64_t nz,
                   double dxsqi, double dysqi, double dzsqi,
                   const double *phi, const double *rho, double *res,
                   int64_t mglevel, const int64_t *bounds, double mgparam,
                   int64_t mgform, int64_t mgform2init,
                   int64_t lcndbndy, int64_t icndbndy, oc_conductors *c);

void oc_restrict3d(int64_t nx, int64_t ny, int64_t nz, const double *res,
                   int64_t nxc, int64_t nyc, int64_t nzc, double *rhocoarse,
                   double ff, const int64_t *bounds);

void oc_expand3d(int64_t nx, int64_t ny, int64_t nz, double *phi,
                 int64_t nxc, int64_t nyc, int64_t nzc, const double *phicoarse,
                 const int64_t *bounds);

int64_t oc_multigrid3dsolve(int64_t nx, int64_t ny, int64_t nz,
                            double dx, double dy, double dz,
                            double *phi, double *rho, const int64_t *bounds,
                            double mgparam, int64_t mgform,
                            int64_t *mgiters, int64_t mgmaxiters,
                            int64_t mgmaxlevels, double *mgerror, double mgtol,
                            int64_t downpasses, int64_t uppasses,
                            int64_t lcndbndy, int64_t icndbndy,
                            oc_conductors *c);

void oc_setrho3d(double *rho, int64_t np, const double *xp, const double *yp,
                 const double *zp, double zgrid, double q, double wght,
                 int64_t nx, int64_t ny, int64_t nz,
                 double dx, double dy, double dz,
                 double xmmin, double ymmin, double zmmin);

void oc_applyrhoboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz,
                                     double *rho, const int64_t *bounds);

void oc_sete3d(const double *phi, int64_t np, const double *xp,
               const double *yp, const double *zp, double zgrid,
               double xmmin, double ymmin, double zmmin,
               double dx, double dy, double dz,
               int64_t nx, int64_t ny, int64_t nz,
               double *ex, double *ey, double *ez);

void oc_epush3d(int64_t np, double *uxp, double *uyp, double *uzp,
                const double *ex, const double *ey, const double *ez,
                double q, double m, double ddt);

void oc_xpush3d(int64_t np, double *xp, double *yp, double *zp,
                const double *uxp, const double *uyp, const double *uzp,
                const double *gaminv, double dtp);

void oc_getzmmnt(int64_t np, const double *xp, const double *yp, const double *zp,
                 const double *uxp, const double *uyp, const double *uzp,
                 const double *gaminv, int64_t usegaminv, double *mm);
void oc_particleboundaries3d(int64_t np, const double *xp, const double *yp,
                             double *zp, double *gaminv, double xmmin,
                             double ymmin, double zmmin, double dx, double dy,
                             double dz, int64_t nx, int64_t ny, int64_t nz,
                             double zgrid, int64_t zperiodic);
int64_t oc_clearpart(int64_t np, int64_t ncol, double **cols, int64_t iuz,
                     int64_t igam, int64_t fillmethod);
void oc_getgrid3d(int64_t np, const double *x, const double *y, const double *z,
                  double *f, int64_t nx, int64_t ny, int64_t nz, const double *grid,
                  double xmin, double xmax, double ymin, double ymax, double zmin,
                  double zmax, int64_t l2symtry, int64_t l4symtry);
void oc_sete_from_e3d(const double *selfe, int64_t np, const double *xp,
                      const double *yp, const double *zp, double zgrid,
                      double xmmin, double ymmin, double zmmin, double dx,
                      double dy, double dz, int64_t nx, int64_t ny, int64_t nz,
                      double *ex, double *ey, double *ez);
void oc_setup_icgrid(int64_t nx, int64_t ny, int64_t nz, int64_t nint,
                     const int64_t *int_indx, int64_t ne, const int64_t *e_indx,
                     int64_t no, const int64_t *o_indx, int32_t *icgrid);
void oc_efield_subgrid(const double *phi, int64_t nx, int64_t ny, int64_t nz, double dx,
                       double dy, double dz, const int32_t *icgrid, int64_t n,
                       const int64_t *indx, const double *dels, const double *volt,
                       double *efield0, double *efieldd);
void oc_fixefield(double *selfe, int64_t nx, int64_t ny, int64_t nz, int64_t nint,
                  const int64_t *int_indx);
void oc_sete3d_conductor(const int32_t *icgrid, const double *e0_even, const double *ed_even,
                         const double *e0_odd, const double *ed_odd, const double *phi,
                         const double *selfe, int64_t np, const double *xp, const double *yp,
                         const double *zp, double zgrid, double xmmin, double ymmin,
                         double zmmin, double dx, double dy, double dz, int64_t nx, int64_t ny,
                         int64_t nz, int64_t efetch, double *ex, double *ey, double *ez);
void oc_getselfe3d(const double *phi, int64_t nx, int64_t ny, int64_t nz,
                   double *selfe, double dx, double dy, double dz,
                   int64_t lzero);

int64_t oc_conductordelfromintercepts(const oc_intercepts *icpt,
                                      oc_conductors *c, double dfill,
                                      double fuzz, int64_t count_only,
                                      int64_t *counts);

void oc_zcylinder_intercepts(double rad, double length, double xcent,
                             double ycent, double zcent, oc_intercepts *icpt,
                             double fuzz, int64_t outside);

void oc_zsrfrv_intercepts(int64_t nn, const double *rsrf, const double *zsrf,
                          double xcent, double ycent, double zcent,
                          oc_intercepts *icpt, int64_t count_only,
                          int64_t *counts);

/* oc_fielddiag.c: line charge and electrostatic energy (dw3d.F). */
void oc_plane_sums(const double *rho, int64_t nx, int64_t ny, int64_t nz,
                   int64_t l2symtry, int64_t l4symtry, double *scrtch);
void oc_linechg_from_planes(const double *scrtch, int64_t nz, double dx, double dy,
                            double dz, double zgrid, double zmmin, double zzmin,
                            double zbeam, double dzz, int64_t nzzarr, double *linechg);
void oc_gtlchg3d(const double *rho, int64_t nx, int64_t ny, int64_t nz, double dx,
                 double dy, double dz, double zgrid, double zmmin, double zzmin,
                 double zbeam, double dzz, int64_t nzzarr, int64_t l2symtry,
                 int64_t l4symtry, double *scrtch, double *linechg);
double oc_getese3d(const double *rho, const double *phi, int64_t nx, int64_t ny, int64_t nz,
                   double dx, double dy, double dz, int64_t l4symtry, int64_t l2symtry);

#ifdef __cplusplus
}
#endif
#endif
