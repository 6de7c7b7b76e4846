/*
 * warpmi.h -- C ABI of libwarpmi.so, the MI355X (gfx950) implementation of
 * WARP's 3-D electrostatic PIC hot path: CIC deposit, multigrid Poisson
 * solve with subgrid-resolved conductors, field gather and leapfrog push.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative code on error;
 *    wm_last_error() then returns a message.  Non-convergence of the
 *    multigrid solve is not an error: it returns 0 with *mgerror > mgtol,
 *    as the reference does.
 *  - Grid arrays are device pointers in the reference's Fortran order
 *    (x fastest): phi is (-1:nx+1,-1:ny+1,-1:nz+1), rho is (0:nx,0:ny,0:nz),
 *    selfe is (0:2,0:nx,0:ny,0:nz).  Particle arrays are device pointers of
 *    length np.  Conductor lists passed to wm_mg3d_set_conductors and the
 *    intercept/conductor set-up routines are host pointers.
 *  - stream is a hipStream_t (NULL = the null stream).  Calls are
 *    asynchronous on that stream except wm_mg3d_solve, which returns after
 *    the solve has finished (mgiters and mgerror are host outputs).
 *  - bounds[6] are (xm, xp, ym, yp, zm, zp) with 0 = Dirichlet,
 *    1 = Neumann, 2 = periodic.
 */
#ifndef WARPMI_H
#define WARPMI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message describing the last error on the calling thread. */
const char *wm_last_error(void);

/* Library and device query: fills name (>= 64 bytes) with the gfx arch of
 * the current device; returns the number of compute units. */
int wm_device_info(char *name, int64_t namelen, int64_t *cus);

/* Measurement only (no reference counterpart): streams n doubles of x
 * through the device, read only (mode 0; out receives nothing for finite
 * data) or write only (mode 1, x = 1.0), with 8-byte lanes.  bench.py runs
 * both over 2 GiB so that a profiler's FETCH_SIZE / WRITE_SIZE counters can
 * be scaled by passes of known size. */
int wm_stream_probe(double *x, double *out, int64_t n, int64_t mode, void *stream);

/* ---------------------------------------------------------------------
 * Multigrid solver (pywarp90/f3d_mgrid.F)
 * ------------------------------------------------------------------- */
typedef struct wm_mg3d wm_mg3d;

/* Creates a solver for an (nx,ny,nz)-cell grid.  The coarsening levels are
 * those of getmglevels (f3d_mgrid.F:3654-3845); mgmaxlevels caps the
 * recursion as in vcycle (f3d_mgrid.F:1293).  Replaces the grid set-up that
 * multigrid3dsolve (f3d_mgrid.F:83) does on every call. */
int wm_mg3d_create(wm_mg3d **out, int64_t nx, int64_t ny, int64_t nz,
                   double dx, double dy, double dz, const int64_t *bounds,
                   int64_t mgmaxlevels);
int wm_mg3d_destroy(wm_mg3d *h);
int wm_mg3d_nlevels(const wm_mg3d *h, int64_t *nlevels);
int wm_mg3d_level_dims(const wm_mg3d *h, int64_t level, int64_t *nxyz,
                       double *dxyz);

/* Conductor lists after checkconductors (f3d_mgrid.F:3964-4326), i.e. the
 * ConductorType of pywarp90/f3d.v:105-164 sorted by level.  Host arrays.
 * int_indx is (n,3), sub_dels/sub_volt are (n,6) in the order
 * mx,px,my,py,mz,pz; *_istart has levels+1 entries (0-based). */
typedef struct {
  int64_t levels;
  int64_t int_n;
  const int64_t *int_indx;
  const double *int_volt;
  const int64_t *int_istart;
  int64_t sub_n[2];
  const int64_t *sub_indx[2];
  const double *sub_dels[2];
  const double *sub_volt[2];
  const int64_t *sub_istart[2];
} wm_conductor_lists;

/* Uploads conductor data into the solver (replaces any previous data).
 * The ConductorType argument of multigrid3dsolve (f3d_mgrid.F:94). */
int wm_mg3d_set_conductors(wm_mg3d *h, const wm_conductor_lists *c);

/* precision: 0 = FP64 V-cycles (the reference's arithmetic),
 *            1 = FP32 V-cycles inside an FP64 defect-correction loop,
 *            2 = as 1 with BF16 storage for the finest-level smoother. */
#define WM_PREC_FP64 0
#define WM_PREC_FP32_DC 1
#define WM_PREC_BF16_DC 2

/* multigrid3dsolve (f3d_mgrid.F:83-500) for iwhich = 0: solves in place on
 * phi (initial guess in, solution with Dirichlet guard cells filled out),
 * scaling rho by mgparam/(eps0*2*(dxsqi+dysqi+dzsqi)) and restoring it as
 * the reference does.  Stops when max|phi_new-phi_old| <= mgtol or after
 * mgmaxiters V-cycles. */
int wm_mg3d_solve(wm_mg3d *h, double *phi, double *rho, double mgparam,
                  int64_t mgform, int64_t mgmaxiters, double mgtol,
                  int64_t downpasses, int64_t uppasses, int64_t lcndbndy,
                  int64_t icndbndy, int64_t precision, int64_t *mgiters,
                  double *mgerror, void *stream);

/* Initial guess of the next solve by linear extrapolation in time:
 * phi <- 2 phi - prev and prev <- the old phi, over n values.  No
 * counterpart in the reference, whose solve starts from the previous phi
 * (fieldsolver.py:1653-1658); the solve still stops on mgtol. */
int wm_phi_extrapolate(double *phi, double *prev, int64_t n, void *stream);

/* No counterpart in the reference (a hint for callers that keep phi
 * between solves, such as a time-step loop): declares that the phi the
 * next wm_mg3d_solve receives is the same array, unchanged, that the last
 * wm_mg3d_solve on h returned.  The next solve then starts from its own
 * red-black copy of that phi instead of converting it again; it does so
 * only for the FP64 reference form (mgform 1) with six Dirichlet bounds, no
 * conductor record on an interior point, an up pass, and after a solve of
 * the same kind that ran at least one V-cycle -- otherwise the hint is
 * ignored.  The result is the same bits either way.  The hint applies to
 * the next solve only.  The per-level entry points (wm_mg3d_sorhalfpass,
 * wm_mg3d_sorpass, wm_mg3d_residual) overwrite the solver's level buffers,
 * so after any of them the next solve converts phi again. */
int wm_mg3d_reuse_phi(wm_mg3d *h);

/* No counterpart in the reference: *reused = 1 when the last wm_mg3d_solve
 * on h started from its own copy of phi (the hint above was honoured),
 * else 0. */
int wm_mg3d_last_reused(const wm_mg3d *h, int64_t *reused);

/* No counterpart in the reference: *level = the first level of the FP64
 * V-cycle's coarse end that the last solve ran as one single-workgroup
 * launch with the levels' arrays in LDS (the levels from there to the
 * coarsest), or -1 when it ran level by level.  The environment variable
 * WARPMI_TAIL_LDS (bytes; 0 switches the single launch off) caps the LDS
 * budget that decides the level, read when the solver is created. */
int wm_mg3d_tail_start(const wm_mg3d *h, int64_t *level);

/* Form of the SOR passes (sorpass3d, f3d_mgrid.F:1949-2093): 0 = two
 * half sweeps, one launch per colour (the default); 1 = both colours'
 * dense updates in one launch on every level where that applies (six
 * Dirichlet bounds, no subgrid record on an interior point, no Neumann
 * subgrid mirror, at most 256 colour elements per row, an even number of
 * passes per V-cycle), reading phi from one red-black buffer and writing
 * the other, with the conductor records in launches of their own before
 * and after it; 2 = that on the finest level only.  Every form gives the
 * same bits.  Changing it drops the captured V-cycles. */
int wm_mg3d_set_sor_form(wm_mg3d *h, int64_t form);

/* Event timing of the finest-level FP64 SOR half sweep (the solver's
 * dominant kernel): while on, wm_mg3d_solve launches its kernels one
 * V-cycle at a time instead of replaying its graph, records HIP events
 * around every finest-level half-sweep launch on its stream and adds the
 * elapsed times up.  Setting it resets the totals. */
int wm_mg3d_set_timing(wm_mg3d *h, int64_t on);
int wm_mg3d_get_timing(const wm_mg3d *h, double *ms_total, int64_t *launches);
/* *fused = 1 when the timed launches were full SOR passes (both colours in
 * one launch, reading one phi buffer and writing another: 3 scalars moved
 * per node) rather than half sweeps (2 scalars per node). */
int wm_mg3d_timing_fused(const wm_mg3d *h, int64_t *fused);

/* Per-kernel entry points on one level, natural-layout device arrays of
 * that level's size.  sorhalfpass3d (f3d_mgrid.F:2095-2291), residual3d
 * (3367-3598), restrict3d (1593-1775, res has 1 guard cell, rhocoarse is
 * the next level's (0:nc)^3), expand3d (1777-1947, phicoarse is the next
 * level's (-1:nc+1)^3). */
int wm_mg3d_sorhalfpass(wm_mg3d *h, int64_t level, int64_t parity,
                        double *phi, const double *rho, double mgparam,
                        int64_t mgform, int64_t lcndbndy, int64_t icndbndy,
                        void *stream);
/* sorpass3d (f3d_mgrid.F:1949-2093): conductor pinning, then each parity's
 * half sweep followed by applyboundaryconditions3d. */
int wm_mg3d_sorpass(wm_mg3d *h, int64_t level, double *phi, const double *rho,
                    double mgparam, int64_t mgform, int64_t lcndbndy,
                    int64_t icndbndy, void *stream);
int wm_mg3d_residual(wm_mg3d *h, int64_t level, const double *phi,
                     const double *rho, double *res, double mgparam,
                     int64_t mgform, int64_t mgform2init, int64_t lcndbndy,
                     int64_t icndbndy, void *stream);
int wm_mg3d_restrict(wm_mg3d *h, int64_t level, const double *res,
                     double *rhocoarse, void *stream);
int wm_mg3d_expand(wm_mg3d *h, int64_t level, double *phi,
                   const double *phicoarse, void *stream);

/* applyboundaryconditions3d (f3d_mgrid.F:3847-3955), ncomp = 1, one guard
 * cell, on a device array (-1:nx+1)^3. */
int wm_applyboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz,
                                 double *u, const int64_t *bounds,
                                 int64_t lwithdirichlet,
                                 int64_t lzerodirichlet, void *stream);

/* ---------------------------------------------------------------------
 * Particles (pywarp90/w3d.F)
 * ------------------------------------------------------------------- */

/* setrho3d -> setrho3ddirect1 (w3d.F:7199-7301, 7757-7963), no symmetry:
 * adds the CIC charge of np particles of charge q and weight wght to rho.
 * Positions must lie in [xmmin, xmmin+nx*dx) etc. (the reference does not
 * check; here out-of-grid particles are skipped and counted in
 * *nskipped, which may be NULL). */
int wm_setrho3d(double *rho, int64_t np, const double *xp, const double *yp,
                const double *zp, double zgrid, double q, double wght,
                int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                double dz, double xmmin, double ymmin, double zmmin,
                int64_t *nskipped, void *stream);

/* Deposit of particles sorted by cell (no reference counterpart as one
 * call; same sums as setrho3d grouped by cell): computes the cell keys,
 * sorts a permutation by key and reduces each cell's charge before adding
 * it to rho.  work must hold wm_setrho3d_sorted_worksize() bytes. */
int64_t wm_setrho3d_sorted_worksize(int64_t np, int64_t nx, int64_t ny,
                                    int64_t nz);
int wm_setrho3d_sorted(double *rho, int64_t np, const double *xp,
                       const double *yp, const double *zp, double zgrid,
                       double q, double wght, int64_t nx, int64_t ny,
                       int64_t nz, double dx, double dy, double dz,
                       double xmmin, double ymmin, double zmmin, void *work,
                       void *stream);

/* applyrhoboundaryconditions3d (w3d.F:2660-2750), serial, no guards. */
int wm_applyrhoboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz,
                                    double *rho, const int64_t *bounds,
                                    void *stream);

/* sete3d -> sete_from_phi_linear3d (w3d.F:3656-3777, 3963-4107): adds E
 * at the particles to ex, ey, ez. */
int wm_sete3d(const double *phi, int64_t np, const double *xp,
              const double *yp, const double *zp, double zgrid,
              double xmmin, double ymmin, double zmmin, double dx, double dy,
              double dz, int64_t nx, int64_t ny, int64_t nz, double *ex,
              double *ey, double *ez, void *stream);

/* epush3d (w3d.F:2752-2777): u += E*q*ddt/m. */
int wm_epush3d(int64_t np, double *uxp, double *uyp, double *uzp,
               const double *ex, const double *ey, const double *ez, double q,
               double m, double ddt, void *stream);

/* xpush3d (w3d.F:3070-3092): x += u*gaminv*dtp. */
int wm_xpush3d(int64_t np, double *xp, double *yp, double *zp,
               const double *uxp, const double *uyp, const double *uzp,
               const double *gaminv, double dtp, void *stream);

/* The "fullv" advance of padvnc3d for B = 0 (w3d.F:1816, 1956-1986):
 * E zeroed and fetched (sete3d), epush3d(dt/2) twice, xpush3d(dt), in one
 * pass.  ex, ey, ez may be NULL; when given they receive E.  nlost (device
 * memory, may be NULL) is incremented by the number of particles whose new
 * position particleboundaries3d would absorb: x <= xmmin or x >= xmmax, the
 * same in y, z-zgrid <= zmmin or z-zgrid >= zmmax (top.F:3550-3560,
 * 3779-3795; the reference passes xmmax = xmmin+nx*dx, ..., w3d.F:11344-11356). */
int wm_gather_push3d(const double *phi, int64_t np, double *xp, double *yp,
                     double *zp, double *uxp, double *uyp, double *uzp,
                     const double *gaminv, double *ex, double *ey,
                     double *ez, double zgrid, double xmmin, double ymmin,
                     double zmmin, double dx, double dy, double dz,
                     int64_t nx, int64_t ny, int64_t nz, double q, double m,
                     double dt, double xmmax, double ymmax, double zmmax,
                     int64_t *nlost, void *stream);

/* padvnc3d(center) for B = 0, no lattice elements (w3d.F:1575-2330), one
 * pass per particle: E zeroed and fetched (sete3d) into ex, ey, ez (may be
 * NULL), then by center
 *   WM_CENTER_FULLV  (w3d.F:1848-1986): epush3d(dt/2) twice, xpush3d(dt);
 *   WM_CENTER_HALFV  (w3d.F:2063-2182): epush3d(dt/2), xpush3d(dt);
 *   WM_CENTER_SYNCHV (w3d.F:2186-2311): epush3d(dt/2), no drift.
 * nlost as in wm_gather_push3d (untouched by synchv, which moves no
 * particle).  wm_gather_push3d is center WM_CENTER_FULLV. */
#define WM_CENTER_FULLV 0
#define WM_CENTER_HALFV 1
#define WM_CENTER_SYNCHV 2
int wm_padvnc3d(const double *phi, int64_t np, double *xp, double *yp, double *zp,
                double *uxp, double *uyp, double *uzp, const double *gaminv, double *ex,
                double *ey, double *ez, double zgrid, double xmmin, double ymmin,
                double zmmin, double dx, double dy, double dz, int64_t nx, int64_t ny,
                int64_t nz, double q, double m, double dt, double xmmax, double ymmax,
                double zmmax, int64_t *nlost, int64_t center, void *stream);

/* wm_padvnc3d with efetch = 3: E interpolated from the precomputed grid
 * field selfe (as wm_getselfe3d leaves it; sete3d -> sete_from_e_linear3d,
 * w3d.F:3702-3715, 4362-4460) instead of the centred differences of phi. */
int wm_padvnc3d_selfe(const double *selfe, int64_t np, double *xp, double *yp, double *zp,
                      double *uxp, double *uyp, double *uzp, const double *gaminv, double *ex,
                      double *ey, double *ez, double zgrid, double xmmin, double ymmin,
                      double zmmin, double dx, double dy, double dz, int64_t nx, int64_t ny,
                      int64_t nz, double q, double m, double dt, double xmmax, double ymmax,
                      double zmmax, int64_t *nlost, int64_t center, void *stream);

/* sete3d with efetch = 3 (w3d.F:3702-3715 -> sete_from_e_linear3d
 * 4362-4460): adds to ex, ey, ez the trilinear interpolation of selfe
 * (3, 0:nx, 0:ny, 0:nz), component fastest, nxguarde = 0.  Particles whose
 * scaled position lies outside [0, n] in any axis get nothing. */
int wm_sete3d_from_e(const double *selfe, int64_t np, const double *xp, const double *yp,
                     const double *zp, double zgrid, double xmmin, double ymmin, double zmmin,
                     double dx, double dy, double dz, int64_t nx, int64_t ny, int64_t nz,
                     double *ex, double *ey, double *ez, void *stream);

/* Conductor-corrected E field (f3d.lcorrectede, pywarp90/f3d.v:152) on the
 * finest level: a handle holding the conductor field data of
 * ConductorType (icgrid, evensubgrid/oddsubgrid efield0 and efieldd,
 * pywarp90/f3d.v:105-164).
 *
 * setupconductorfielddata (f3d_mgrid.F:4581-4633): classifies the nodes
 * (0:nx, 0:ny, 0:nz) from the level-0 part of the lists (interior nodes,
 * then the even, then the odd subgrid points; a later list overwrites an
 * earlier one).  Host work plus uploads; call again when the conductors
 * change. */
typedef struct wm_condfield wm_condfield;
int wm_condfield_create(wm_condfield **out, int64_t nx, int64_t ny, int64_t nz,
                        const wm_conductor_lists *c);
int wm_condfield_destroy(wm_condfield *cf);

/* getefieldatconductorsubgrid (f3d_mgrid.F:4635-4735): the one-sided E at
 * every level-0 subgrid point from phi (-1:nx+1)^3 (dx, dy, dz of the
 * finest level), and its extrapolation into the conductor where the
 * neighbour is an interior node.  Does nothing when either parity list is
 * empty, as the reference. */
int wm_condfield_efield(wm_condfield *cf, const double *phi, double dx, double dy,
                        double dz, void *stream);

/* fixefieldatconductorpoints (f3d_mgrid.F:4737-4772): doubles selfe
 * (3, 0:nx, 0:ny, 0:nz) at the level-0 interior nodes (once per listing). */
int wm_condfield_fixefield(const wm_condfield *cf, double *selfe, void *stream);

/* The handle's arrays, for inspection: nsub[2] gets the record counts of
 * the even and odd lists; each non-NULL destination (device or host
 * memory) gets a copy of icgrid (int32, (nx+1)(ny+1)(nz+1), x fastest: 0
 * vacuum, 2147483647 interior, +(r+1) even record r, -(r+1) odd record
 * r), efield0 (3 per record) and efieldd (6 per record) of the even and
 * odd lists.  Returns after the copies are done. */
int wm_condfield_read(const wm_condfield *cf, int64_t *nsub, int32_t *icgrid,
                      double *e0_even, double *ed_even, double *e0_odd, double *ed_odd,
                      void *stream);

/* sete3dwithconductor (f3d_mgrid.F:4774-5010 with GETEEFROME 5164-5233),
 * efetch 1 (from phi) or 3 (from selfe, as wm_getselfe3d and
 * wm_condfield_fixefield leave it), no symmetries: adds E to ex, ey, ez.
 * Cells with a conductor node take the weighted one-sided fields;
 * particles outside [0, n] on some axis get nothing. */
int wm_sete3d_conductor(const wm_condfield *cf, const double *phi, const double *selfe,
                        int64_t efetch, int64_t np, const double *xp, const double *yp,
                        const double *zp, double zgrid, double xmmin, double ymmin,
                        double zmmin, double dx, double dy, double dz, double *ex,
                        double *ey, double *ez, void *stream);

/* wm_padvnc3d with E from sete3dwithconductor (efetch 1 or 3) instead of
 * sete3d: the advance padvnc3d makes under f3d.lcorrectede
 * (w3d.F:1575-2330 -> fetche3d -> multigrid.py:528-546). */
int wm_padvnc3d_conductor(const wm_condfield *cf, const double *phi, const double *selfe,
                          int64_t efetch, int64_t np, double *xp, double *yp, double *zp,
                          double *uxp, double *uyp, double *uzp, const double *gaminv,
                          double *ex, double *ey, double *ez, double zgrid, double xmmin,
                          double ymmin, double zmmin, double dx, double dy, double dz,
                          double q, double m, double dt, double xmmax, double ymmax,
                          double zmmax, int64_t *nlost, int64_t center, void *stream);

/* getzmmnt (top.F:521-1360), itask 2: the whole-beam moment sums
 * zmmnts0(1:33) of np particles into mm (device, 33 doubles, overwritten;
 * mm[k-1] = zmmnts0(k)), summed per block then over blocks in a fixed
 * order.  usegaminv selects the lrelativ branch (velocities u*gaminv).
 * work must hold wm_getzmmnt_worksize(np) bytes. */
int64_t wm_getzmmnt_worksize(int64_t np);
int wm_getzmmnt(int64_t np, const double *xp, const double *yp,
                const double *zp, const double *uxp, const double *uyp,
                const double *uzp, const double *gaminv, int64_t usegaminv,
                double *mm, void *work, void *stream);

/* gtlchg3dfromrho (dw3d.F:229-320; called by MultiGrid3D.gtlchg,
 * multigrid.py:433-449) for one processor: planes[iz] (device, nz+1
 * doubles) receives the charge summed over z plane iz of rho ((nx+1)(ny+1)
 * (nz+1) doubles, no guard cells) -- the rows iy >= 1 doubled under
 * l2symtry, the quadrant weights 1/2/4 under l4symtry (l2symtry wins when
 * both are set, as in the reference) -- and linechg[0..nzzarr] (device) the
 * line charge [C/m] interpolated onto the Z_arrays mesh zzmin + iz*dzz of
 * the beam frame (zbeam).  Each plane is summed in the reference's order;
 * the results are the reference's bits. */
int wm_gtlchg3d(const double *rho, int64_t nx, int64_t ny, int64_t nz,
                double dx, double dy, double dz, double zgrid, double zmmin,
                double zzmin, double zbeam, double dzz, int64_t nzzarr,
                int64_t l2symtry, int64_t l4symtry, double *planes,
                double *linechg, void *stream);

/* getese3dfromrhophi (dw3d.F:708-768; called by MultiGrid3D.getese,
 * multigrid.py:451-467): ese[0] (device) = dx*dy*dz/2 * sum of rho*phi
 * over the mesh nodes, rho without guard cells, phi with one guard cell,
 * with the reference's symmetry factors (l4symtry wins when both are set)
 * and summation order (columns over z, then over y, then over x).  work
 * must hold wm_getese3d_worksize(nx, ny) bytes. */
int64_t wm_getese3d_worksize(int64_t nx, int64_t ny);
int wm_getese3d(const double *rho, const double *phi, int64_t nx, int64_t ny,
                int64_t nz, double dx, double dy, double dz, int64_t l4symtry,
                int64_t l2symtry, void *work, double *ese, void *stream);

/* No counterpart in the reference: wm_getese3d for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The column sums live in the library's own device
 * scratch (1 MiB per GPU, held until the process ends) instead of a work
 * argument, so (nx+1)*(ny+1) may be at most 131072; larger meshes return
 * an error.  Same result, bit for bit.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (other streams, other host
 * threads) race. */
int wm_getese3d_noalloc(const double *rho, const double *phi, int64_t nx, int64_t ny,
                        int64_t nz, double dx, double dy, double dz, int64_t l4symtry,
                        int64_t l2symtry, double *ese, void *stream);

/* particleboundaries3d (w3d.F:11292-11450) for absorbing boundaries on
 * all six faces (pboundxy = pbound0 = pboundnz = absorb), followed by
 * processlostpart -> clearpart (top.F:4936-5070, 4730-4893) on one species
 * block of np particles.  cols (a host array of ncol device pointers, 7 <=
 * ncol <= 16) starts xp, yp, zp, uxp, uyp, uzp, gaminv; further columns
 * (ex, ey, ez, ...) move along.  A particle is flagged lost (gaminv = 0) when
 * x <= xmmin or x >= xmmin+nx*dx, the same in y, or z-zgrid >= zmmin+nz*dz
 * or z-zgrid <= zmmin (xparticleboundaries top.F:3550-3560, its y twin,
 * zparticleboundaries top.F:3779-3795).  With zbound = 2 (pbound0 = pboundnz
 * = periodic) z wraps by whole periods instead (top.F:3762-3778) and only x
 * and y absorb; zbound = 0 absorbs in z.  Then every particle with gaminv
 * <= 0 is removed with clearpart's fillmethod (top.v clearlostpart):
 *   1: in place, holes filled with live particles taken from the end (the
 *      reference's default; dstcols NULL);
 *   2: live particles shifted down, order kept; written to dstcols (ncol
 *      other device columns), the input columns are left as they were.
 * *nps_out (device memory) receives the new particle count.  work (device)
 * must hold wm_clearpart_worksize(np) bytes. */
int64_t wm_clearpart_worksize(int64_t np);
int wm_particleboundaries3d(int64_t np, int64_t ncol, double *const *cols, double xmmin,
                            double ymmin, double zmmin, double dx, double dy, double dz,
                            int64_t nx, int64_t ny, int64_t nz, double zgrid, int64_t zbound,
                            int64_t fillmethod, double *const *dstcols, int64_t *nps_out,
                            void *work, void *stream);

/* clearpart (top.F:4730-4893) alone: removes the particles already flagged
 * with gaminv <= 0; arguments as wm_particleboundaries3d. */
int wm_clearpart(int64_t np, int64_t ncol, double *const *cols, int64_t fillmethod,
                 double *const *dstcols, int64_t *nps_out, void *work, void *stream);

/* getgrid3d (dtop.F:2818-2920) over a node-centred grid (0:nx,0:ny,0:nz)
 * spanning [xmin,xmax]x[ymin,ymax]x[zmin,zmax]: f at each position inside
 * the grid is the trilinear interpolation, in the reference's order of
 * operations; f is left unchanged outside.  With l4symtry x and y, with
 * l2symtry y, are reflected about 0 first.  MultiGrid3D.fetchphi calls it
 * with phi including its guard cells (multigrid.py:560-575): nx+2 cells
 * from xmmin-dx to xmmax+dx, ... */
int wm_getgrid3d(int64_t np, const double *x, const double *y, const double *z,
                 double *f, int64_t nx, int64_t ny, int64_t nz, const double *grid,
                 double xmin, double xmax, double ymin, double ymax, double zmin,
                 double zmax, int64_t l2symtry, int64_t l4symtry, void *stream);

/* particlesortxyzwithcopy (top.F:5095-5138) with particlesortbyindex
 * (top.F:5140-5203): the permutation that orders np particles by the block
 * index int(|x-xmmin|/dx) + (nx+1)*(iy + (ny+1)*iz), stable, so equal
 * blocks keep their order as in the reference's counting sort.  perm[i]
 * (device, np entries) is the old index of the particle that goes to
 * position i.  work must hold wm_particle_sort_worksize() bytes.  Indices
 * outside [0, n] are clamped (the reference requires in-grid particles). */
int64_t wm_particle_sort_worksize(int64_t np, int64_t nx, int64_t ny,
                                  int64_t nz);
int wm_particle_sort(int64_t np, const double *xp, const double *yp,
                     const double *zp, double xmmin, double ymmin,
                     double zmmin, double dx, double dy, double dz,
                     int64_t nx, int64_t ny, int64_t nz, uint32_t *perm,
                     void *work, void *stream);

/* particlecopysorteddata (top.F:5205-5222) for up to 16 columns at once:
 * dst[c][i] = src[c][perm[i]].  src and dst are host arrays of ncol device
 * pointers; dst columns must not alias src columns. */
int wm_permute_columns(int64_t np, int64_t ncol, const double *const *src,
                       double *const *dst, const uint32_t *perm,
                       void *stream);

/* getselfe3d (w3d.F:6023-6109), nxguarde = 0. */
int wm_getselfe3d(const double *phi, int64_t nx, int64_t ny, int64_t nz,
                  double *selfe, double dx, double dy, double dz,
                  int64_t lzero, void *stream);

/* ---------------------------------------------------------------------
 * 3-D U-Net warm start (no reference counterpart: the learned initial
 * guess of the multigrid solve).  Activations are channels-last bf16
 * (uint16 bit patterns), channels multiples of 16; weights are fp32
 * masters w[co][ci][27] packed by wm_unet_pack into the forward layout
 * wp and the mirrored/transposed layout wt of the backward-over-input
 * convolution.
 * ------------------------------------------------------------------- */

/* 3x3x3 convolution, zero padding 1, bf16 MFMA (K = 32) with fp32
 * accumulation, input tiles staged in LDS:
 * y = leaky_relu(conv(x) + bias, slope) (slope 1: identity; bias may be
 * NULL).  Output to y (bf16) and/or yf (fp32), either may be NULL.
 * cin and cout multiples of 16, cout <= 128. */
int wm_unet_conv3(const uint16_t *x, int64_t cin, uint16_t *y, float *yf,
                  int64_t cout, int64_t D, int64_t H, int64_t W,
                  const uint16_t *wp, const float *bias, double slope,
                  void *stream);
/* dw[co][ci][tap] += sum_v x[v+tap][ci] g[v][co]; db[co] += sum_v g[v][co]
 * (db may be NULL). */
int wm_unet_conv3_wgrad(const uint16_t *x, int64_t cin, const uint16_t *g,
                        int64_t cout, int64_t D, int64_t H, int64_t W,
                        float *dw, float *db, void *stream);
/* g *= leaky_relu'(y) elementwise (n values). */
int wm_unet_act_bwd(uint16_t *g, const uint16_t *y, double slope, int64_t n,
                    void *stream);
/* Packs the fp32 masters w[co][ci][27] as bf16 into wp (forward) and wt
 * (backward over the input: taps mirrored, channel roles exchanged), each
 * of wm_unet_pack_size(cin, cout) = 30*cin*cout values: per 16-channel
 * block, z offset of the tap and pair of its 9 (y, x) taps (a zero tenth
 * tap pads the last pair) the K = 32 operand of one MFMA step. */
int64_t wm_unet_pack_size(int64_t cin, int64_t cout);
int wm_unet_pack(const float *w, int64_t cin, int64_t cout, uint16_t *wp,
                 uint16_t *wt, void *stream);
/* 2x2x2 average pooling of x (D,H,W,c) into y; backward: x is the pooled
 * gradient, y the full-size gradient. */
int wm_unet_pool(const uint16_t *x, int64_t D, int64_t H, int64_t W, int64_t c,
                 uint16_t *y, int64_t backward, void *stream);
/* Nearest x2 upsampling of x (D,H,W,cx) into channels [0,cx) of y (cy
 * channels); backward sums the eight children into x-sized y. */
int wm_unet_up(const uint16_t *x, int64_t D, int64_t H, int64_t W, int64_t cx,
               uint16_t *y, int64_t cy, int64_t backward, void *stream);
/* The decoder's concatenated input: y (2D,2H,2W, cx+cs) = [nearest x2
 * upsampling of x (D,H,W,cx) | skip (2D,2H,2W,cs)]. */
int wm_unet_upcat(const uint16_t *x, int64_t D, int64_t H, int64_t W, int64_t cx,
                  const uint16_t *skip, int64_t cs, uint16_t *y, void *stream);
int wm_unet_chcopy(const uint16_t *src, int64_t sc, int64_t soff,
                   uint16_t *dst, int64_t dc, int64_t doff, int64_t c,
                   int64_t nvox, void *stream);
/* a += b, n bf16 values. */
int wm_unet_add(uint16_t *a, const uint16_t *b, int64_t n, void *stream);
/* Adam step (bias-corrected) on n fp32 weights. */
int wm_unet_adam(float *w, const float *g, float *m1, float *m2, int64_t n,
                 double lr, double b1, double b2, double eps, int64_t step,
                 void *stream);
/* Network input (rho*srho, mask, phi*sphi, zeros) over nodes 0..n-1 of
 * FP64 natural rho (0:n)^3 and phi (-1:n+1)^3; mask is one byte per node
 * (may be NULL). */
int wm_unet_input(const double *rho, const double *phi, const uint8_t *mask,
                  int64_t nx, int64_t ny, int64_t nz, double srho, double sphi,
                  uint16_t *x, void *stream);
/* phi(0..n-1) += out[.][0]/sphi where mask is 0. */
int wm_unet_apply(const float *out, const uint8_t *mask, int64_t nx, int64_t ny,
                  int64_t nz, double sphi, double *phi, void *stream);
/* Mean-square loss of channel 0 of out (16-channel fp32) against tgt:
 * *loss += mean, g (16-channel bf16) = its gradient. */
int wm_unet_mse(const float *out, const float *tgt, int64_t n, uint16_t *g,
                float *loss, void *stream);

/* The learned initial guess at half resolution (no counterpart in the
 * reference; it replaces the previous-phi guess of fieldsolver.py:1653-1658
 * with the linear extrapolation in time plus the network's estimate of its
 * error).  A coarse voxel is the 2x2x2 block of mesh nodes 2i..2i+1 of the
 * nodes 0..n-1 (nx, ny, nz even); coarse arrays are fp32 (nz/2, ny/2, nx/2).
 * wm_unet_half_mean: block means of an FP64 natural node array with guard
 * (0 or 1) guard cells.  wm_unet_second_diff: d2 = r0 - 2 r1 + r2 and
 * atomic maxima mx[0] = max|d2|, mx[1] = max|r0| (float bits, mx zeroed
 * beforehand).  wm_unet_learn_input: the network input d2/mx[0], maskh,
 * r0/mx[1].  wm_unet_learn_target: block means of p0 - 2 p1 + p2 (FP64
 * phi arrays with guard cells) over cphys*mx[0].  wm_unet_learn_apply:
 * phi += cphys*mx[0] * (trilinear value of out's channel 0) on nodes
 * 1..n-1 where mask (one byte per node 0..n-1, may be NULL) is 0. */
int wm_unet_half_mean(const double *src, int64_t nx, int64_t ny, int64_t nz,
                      int64_t guard, float *dst, void *stream);
int wm_unet_second_diff(const float *r0, const float *r1, const float *r2, int64_t n,
                        float *d2, uint32_t *mx, void *stream);
int wm_unet_learn_input(const float *d2, const float *r0, const float *maskh,
                        const uint32_t *mx, int64_t n, uint16_t *x, void *stream);
int wm_unet_learn_target(const double *p0, const double *p1, const double *p2,
                         int64_t nx, int64_t ny, int64_t nz, const uint32_t *mx,
                         double cphys, float *tgt, void *stream);
int wm_unet_learn_apply(const float *out, const uint32_t *mx, double cphys, int64_t nx,
                        int64_t ny, int64_t nz, const uint8_t *mask, double *phi,
                        void *stream);

/* ---------------------------------------------------------------------
 * Conductor set-up on the host (pywarp90/f3d_conductors.F)
 * ------------------------------------------------------------------- */

/* Grid-line intercepts of one object on one MG level
 * (ConductorInterceptType).  Host arrays: xintercepts (0:nz,0:ny,nxicpt),
 * yintercepts (0:nz,0:nx,nyicpt), zintercepts (0:ny,0:nx,nzicpt) with the
 * intercept index fastest; voltages/condids alike. */
typedef struct {
  double xmmin, ymmin, zmmin, dx, dy, dz;
  int64_t nx, ny, nz, ix, iy, iz, mglevel;
  int64_t nxicpt, nyicpt, nzicpt;
  double *xintercepts, *yintercepts, *zintercepts;
  double *xvoltages, *yvoltages, *zvoltages;
  int64_t *xcondids, *ycondids, *zcondids;
} wm_intercepts;

/* ZCylinderConductorFnew (f3d_conductors.F:2496-2583, outside = 0) and
 * ZCylinderOutConductorFnew (3158-3256, outside = 1). */
int wm_zcylinder_intercepts(double rad, double length, double xcent,
                            double ycent, double zcent, wm_intercepts *icpt,
                            double fuzz, int64_t outside);

/* ZSrfrvConductorFnew (f3d_conductors.F:7040-7855) for outlines made of
 * straight segments (ZAnnulus and polygonal ZSrfrvIn). */
int wm_zsrfrv_intercepts(int64_t nn, const double *rsrf, const double *zsrf,
                         double xcent, double ycent, double zcent,
                         wm_intercepts *icpt);

/* Growable conductor lists owned by the library (host memory). */
typedef struct wm_condlist wm_condlist;
int wm_condlist_create(wm_condlist **out);
int wm_condlist_destroy(wm_condlist *c);
/* conductordelfromintercepts (f3d_conductors.F:12-728) appending to c. */
int wm_conductordelfromintercepts(const wm_intercepts *icpt, wm_condlist *c,
                                  double dfill, double fuzz);
/* checkconductors (f3d_mgrid.F:3964-4326): levels, clean-up, sort. */
int wm_checkconductors(wm_condlist *c, int64_t nx, int64_t ny, int64_t nz,
                       double dx, double dy, double dz);
/* Read access to the lists (pointers valid until the next modification). */
int wm_condlist_view(const wm_condlist *c, wm_conductor_lists *view);

#ifdef __cplusplus
}
#endif
#endif
