// Multigrid Poisson solver with subgrid-resolved conductors for gfx950:
// the MI355X implementation of multigrid3dsolve (pywarp90/f3d_mgrid.F).
//
// The host side builds the level hierarchy (getmglevels), the transfer
// tables (restrict3d/expand3d weights), the conductor stencil records and
// masks, and drives the V-cycle.  One V-cycle plus its convergence test is
// captured once as a HIP graph and replayed; the convergence kernel sets a
// device flag that turns every later kernel into a no-op, so several
// V-cycles can be enqueued before the host looks at the result.
#include "wm_internal.h"
#include "mg3d_kernels.h"
#include "mg3d_fused.h"
#include "mg3d_tail.h"
#include "../../include/warpmi.h"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

namespace {

enum { MODE_RAW = SIDE_RAW, MODE_MUL = SIDE_MUL, MODE_DIV = SIDE_DIV, MODE_CONST = SIDE_CONST };

struct Sym {
  int mode = MODE_RAW;
  double k = 0.;
};

// Row kernels are instantiated for KW = ceil(W / 64) groups of 64 elements
// per lane pass (1-4, else 8 with a loop over the row).
static int kw_of(int W) {
  int k = (W + 63) / 64;
  return k <= 4 ? k : 8;
}
#define WM_KW_DISPATCH(W, LAUNCH)                 \
  switch (kw_of(W)) {                             \
    case 1: { constexpr int KW = 1; LAUNCH; } break; \
    case 2: { constexpr int KW = 2; LAUNCH; } break; \
    case 3: { constexpr int KW = 3; LAUNCH; } break; \
    case 4: { constexpr int KW = 4; LAUNCH; } break; \
    default: { constexpr int KW = 8; LAUNCH; } break; \
  }

static Sym sym_mul(Sym a, double k) {
  if (a.mode == MODE_RAW) return Sym{MODE_MUL, k};
  if (a.mode == MODE_CONST) return Sym{MODE_CONST, a.k * k};
  throw wm_error("subgrid stencil: unexpected double scaling of a neighbour");
}
static Sym sym_div(Sym a, double k) {
  if (a.mode == MODE_RAW) return Sym{MODE_DIV, k};
  if (a.mode == MODE_CONST) return Sym{MODE_CONST, a.k / k};
  throw wm_error("subgrid stencil: unexpected double scaling of a neighbour");
}

// handlesubgrid2 (f3d_mgrid.F:2628-2764) with the neighbour values kept
// symbolic; every number is produced by the reference's own expression.
static void handlesubgrid2(int im, int ip, Sym &pm, Sym &pp, double dsqicos,
                           const double *d, const double *v, double voltfac,
                           double &denom, double &ppp) {
  double adel;
  if (0. < d[im] && d[im] < +1. && 0. < d[ip] && d[ip] < +1.) {
    pm = Sym{MODE_CONST, voltfac * v[im] * 2. / (d[im] * (d[ip] + d[im]))};
    pp = Sym{MODE_CONST, voltfac * v[ip] * 2. / (d[ip] * (d[ip] + d[im]))};
    denom = denom + 2. * (1. - d[ip] * d[im]) / (d[ip] * d[im]) * dsqicos;
    ppp = std::min(ppp, d[im]);
    ppp = std::min(ppp, d[ip]);
  } else if (0. >= d[im] && d[im] > -1. && 0. >= d[ip] && d[ip] > -1.) {
    pm = Sym{MODE_CONST, 0.};
    pp = Sym{MODE_CONST, 0.};
    denom = denom - 2. * dsqicos;
    if (std::fabs(d[im]) > 0.) ppp = std::min(ppp, std::fabs(d[im]));
    else ppp = std::min(ppp, 1. - 1.e-9);
    if (std::fabs(d[ip]) > 0.) ppp = std::min(ppp, std::fabs(d[ip]));
    else ppp = std::min(ppp, 1. - 1.e-9);
  } else if (0. < d[im] && d[im] < +1. && 0. >= d[ip] && d[ip] > -1.) {
    adel = std::fabs(d[ip]);
    pm = Sym{MODE_CONST, voltfac * v[im] / (d[im] * (adel + 0.5 * d[im]))};
    pp = Sym{MODE_CONST, 0.};
    denom = denom + (1. - 2. * d[im] * adel - d[im]) / (d[im] * (adel + 0.5 * d[im])) * dsqicos;
    ppp = std::min(ppp, d[im]);
    if (adel > 0.) ppp = std::min(ppp, adel);
    else ppp = std::min(ppp, 1. - 1.e-9);
  } else if (0. >= d[im] && d[im] > -1. && 0. < d[ip] && d[ip] < +1.) {
    adel = std::fabs(d[im]);
    pm = Sym{MODE_CONST, 0.};
    pp = Sym{MODE_CONST, voltfac * v[ip] / (d[ip] * (adel + 0.5 * d[ip]))};
    denom = denom + (1. - 2. * d[ip] * adel - d[ip]) / (d[ip] * (adel + 0.5 * d[ip])) * dsqicos;
    if (adel > 0.) ppp = std::min(ppp, adel);
    else ppp = std::min(ppp, 1. - 1.e-9);
    ppp = std::min(ppp, d[ip]);
  } else {
    if (0. < d[im] && d[im] < +1.) {
      pm = Sym{MODE_CONST, voltfac * v[im] * 2. / (d[im] * (1. + d[im]))};
      pp = sym_mul(pp, 2. / (1. + d[im]));
      denom = denom + 2. * (1. - d[im]) / d[im] * dsqicos;
      ppp = std::min(ppp, d[im]);
    } else if (0. >= d[im] && d[im] > -1.) {
      adel = std::fabs(d[im]);
      pm = Sym{MODE_CONST, 0.};
      pp = sym_div(pp, 0.5 + adel);
      denom = denom - 2. * adel / (adel + 0.5) * dsqicos;
      if (adel > 0.) ppp = std::min(ppp, adel);
      else ppp = std::min(ppp, 1. - 1.e-9);
    }
    if (0. < d[ip] && d[ip] < +1.) {
      pp = Sym{MODE_CONST, voltfac * v[ip] * 2. / (d[ip] * (1. + d[ip]))};
      pm = sym_mul(pm, 2. / (1. + d[ip]));
      denom = denom + 2. * (1. - d[ip]) / d[ip] * dsqicos;
      ppp = std::min(ppp, d[ip]);
    } else if (0. >= d[ip] && d[ip] > -1.) {
      adel = std::fabs(d[ip]);
      pp = Sym{MODE_CONST, 0.};
      pm = sym_div(pm, 0.5 + adel);
      denom = denom - 2. * adel / (adel + 0.5) * dsqicos;
      if (adel > 0.) ppp = std::min(ppp, adel);
      else ppp = std::min(ppp, 1. - 1.e-9);
    }
  }
}

// Host image of one colour's subgrid records.
struct SubHost {
  std::vector<int> pos;
  std::vector<uint32_t> modes, zero;
  std::vector<double> k, denom, fac;
  int n() const { return (int)modes.size(); }
};

template <typename T>
struct SubDev {
  DevBuf<int> pos;
  DevBuf<uint32_t> modes, zero;
  DevBuf<T> k, denom, fac;
  int n = 0;
  bool anyzero = false;  // some record zeroes a neighbour's residual
  void upload(const SubHost &h) {
    n = h.n();
    anyzero = false;
    for (uint32_t z : h.zero) anyzero = anyzero || z != 0;
    pos.upload(h.pos);
    modes.upload(h.modes);
    zero.upload(h.zero);
    std::vector<T> kk(h.k.begin(), h.k.end()), dd(h.denom.begin(), h.denom.end()),
        ff(h.fac.begin(), h.fac.end());
    k.upload(kk);
    denom.upload(dd);
    fac.upload(ff);
  }
  SubList<T> view() const {
    return SubList<T>{n, pos.p, modes.p, k.p, denom.p, fac.p, zero.p};
  }
};

struct RecKey {
  double mgparam = -1.;
  int mgform = -1, icndbndy = -1, lcndbndy = -1, mgform2init = -1;
  bool operator==(const RecKey &o) const {
    return mgparam == o.mgparam && mgform == o.mgform && icndbndy == o.icndbndy &&
           lcndbndy == o.lcndbndy && mgform2init == o.mgform2init;
  }
};

// Conductor records of one level for one precision.
template <typename T>
struct CondDev {
  RecKey sorkey, reskey[2];
  SubDev<T> sor[2];
  SubDev<T> res[2][2];  // [mgform2init][colour]
  // both colours' residual records in natural order (z, y, x): entry
  // (colour << 31) | index, so that neighbouring threads of k_residual_rec2
  // read neighbouring points of phi
  DevBuf<uint32_t> resord[2];
  void reset() {
    sorkey = reskey[0] = reskey[1] = RecKey();
    for (int p = 0; p < 2; p++) {
      sor[p].upload(SubHost());
      res[0][p].upload(SubHost());
      res[1][p].upload(SubHost());
      resord[p].release();
    }
  }
};

struct Level {
  int nx, ny, nz;
  double dx, dy, dz, dxsqi, dysqi, dzsqi;
  RBGeom g;
  // buffers (allocated on demand)
  DevBuf<double> phi64, rho64, res64;
  DevBuf<float> phi32, rho32, res32;
  DevBuf<bf16> phi16, rho16;  // finest level of the BF16 mode
  // the other buffer of each phi for the fused SOR pass, which reads one and
  // writes the other; after a V-cycle (an even number of passes on every
  // level) phi is back in the first.  Boundary values match the first's.
  DevBuf<double> phi64b;
  DevBuf<float> phi32b;
  DevBuf<bf16> phi16b;
  // transfer to the next coarser level
  bool coarser = false;
  double ff = 0.;
  DevBuf<int> rlo[3], rhi[3];
  DevBuf<double> rw[3];
  int rcmin[3], rcmax[3];
  // widest window and the fine span of the computed range, per axis
  int rwidth[3] = {0, 0, 0}, rflo[3] = {0, 0, 0}, rfhi[3] = {-1, -1, -1};
  // widest fine span of a tile of 32 (x) or 8 (y) consecutive coarse nodes
  int rspan32[3] = {0, 0, 0}, rspan8[3] = {0, 0, 0};
  DevBuf<int> ej[3];
  DevBuf<double> ew[3];
  int elo[3], ehi[3];
  bool rstd[3] = {false, false, false};  // restriction windows are the plain 2:1 ones
  // conductors
  int nint = 0;
  DevBuf<int> int_pos;
  DevBuf<double> int_volt;
  DevBuf<uint32_t> mask;
  DevBuf<uint32_t> imask;  // interior conductor points only (zeroed by k_expand_rb)
  bool hasmask = false;
  // an active subgrid record sits on an interior (pinned) point, so the
  // pin after each half sweep is not a no-op
  bool sub_on_interior = false;
  std::vector<int> subpos_h[2];
  std::vector<double> subdels_h[2], subvolt_h[2];
  DevBuf<int> subpos_d[2];
  DevBuf<double> subdels_d[2];
  bool neumann = false;
  CondDev<double> c64;
  CondDev<float> c32;
  RecKey maskkey;
};

}  // namespace

struct wm_mg3d {
  int64_t nx, ny, nz;
  double dx, dy, dz;
  int bounds[6];
  std::vector<std::unique_ptr<Level>> lv;
  // level-0 extras
  DevBuf<double> save64, rhosave64;
  DevBuf<double> convparts;  // per-wavefront maxima of the fused convergence test
  // The last FP64 solve left its phi in phi64 with save64 equal to it (or,
  // last_pp, equal at every point no sweep writes) and
  // the pinned points at their values (last_ok); when the caller declares
  // its phi unchanged since (wm_mg3d_reuse_phi: reuse_next), the next solve
  // starts from that state.
  bool last_ok = false, reuse_next = false;
  bool last_pp = false;  // the last solve exchanged the phi and save buffers
  bool last_reused = false;  // the last solve started from that state
  const double *last_phi = nullptr;
  DevBuf<SolveState> state;
  DevBuf<double> partials;  // per-block maxima of the convergence test
  SolveState *state_h = nullptr;  // pinned
  hipStream_t s = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // Which of the two level-0 FP64 buffers holds phi: the V-cycles that
  // keep their starting phi in the save buffer exchange the two (phi64 and
  // save64 swap their storage, orient flips).
  int orient = 0;
  // captured V-cycle graphs, per orientation at their start
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraphExec_t gexec_rest[2] = {nullptr, nullptr};
  // n V-cycles of the gexec_rest form in one graph, keyed by (orientation
  // at its start, n): one launch for the V-cycles a solve usually needs
  std::map<std::pair<int, int>, hipGraphExec_t> gmulti;
  void drop_graphs() {
    for (int o = 0; o < 2; o++) {
      if (gexec[o]) (void)hipGraphExecDestroy(gexec[o]);
      if (gexec_rest[o]) (void)hipGraphExecDestroy(gexec_rest[o]);
      gexec[o] = gexec_rest[o] = nullptr;
    }
    for (auto &e : gmulti) (void)hipGraphExecDestroy(e.second);
    gmulti.clear();
  }
  struct GKey {
    int prec = -1, mgform = -1, down = -1, up = -1, lcnd = -1, icnd = -1, maxit = -1;
    double mgparam = -1., mgtol = -1.;
    bool operator==(const GKey &o) const {
      return prec == o.prec && mgform == o.mgform && down == o.down && up == o.up &&
             lcnd == o.lcnd && icnd == o.icnd && maxit == o.maxit && mgparam == o.mgparam &&
             mgtol == o.mgtol;
    }
  } gkey;
  int last_iters = 0;
  // SOR pass form: 0 two half sweeps, 1 the fused two-colour pass where it
  // applies, 2 the fused pass on the finest level only
  // (wm_mg3d_set_sor_form; WARPMI_SOR_FUSED=1 or 2 sets the default)
  int sor_form = getenv("WARPMI_SOR_FUSED") ? std::min(2, std::max(0, atoi(getenv("WARPMI_SOR_FUSED")))) : 0;
  bool conductors_set = false;
  // The coarse end of the FP64 V-cycle in one launch (k_vtail): levels
  // start..nlevels-1, described by the device table tailtab (fixed address,
  // contents rewritten between solves when they change).  The levels whose
  // arrays fit in tail_lds bytes of LDS qualify (WARPMI_TAIL_LDS sets a
  // smaller budget; 0 switches the tail off).
  struct Tail {
    bool on = false;
    int start = -1;
    TailRun run{};
    std::vector<TailLevel> host;  // contents of tailtab
    std::vector<TailItem> items;  // contents of tailitems
    std::vector<TailOp> ops;      // contents of tailops
    int down = -1, up = -1;       // the pass counts the steps are for
  } tail;
  DevBuf<TailLevel> tailtab;
  DevBuf<TailItem> tailitems;
  DevBuf<TailOp> tailops;
  int dc_double_from = -1;  // levels from here on run in FP64 in the FP32/BF16 V-cycles
  bool tail_used = false;  // the last solve was FP64 and ran the tail
  size_t tail_lds = 0;
  int ncu = 256;                 // compute units of the device
  size_t lds_per_cu = 160 * 1024;
  // Event timing of the finest-level SOR half sweeps (eager mode only).
  struct Prof {
    bool on = false;
    std::vector<hipEvent_t> ev;
    int used = 0;
    double ms = 0.;
    long long count = 0;
    bool fused = false;  // the timed launches were full fused passes
  } prof;
  ~wm_mg3d() {
    for (hipEvent_t e : prof.ev) (void)hipEventDestroy(e);
    drop_graphs();
    if (state_h) (void)hipHostFree(state_h);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

static int blocks(long long n) { return (int)std::max<long long>(1, cdiv(n, 256)); }

// The natural-order index of a level's residual records, for lists long
// enough that the locality pays for the extra indirection (measured: 55 ->
// 47 us for 324k records; a few thousand records are latency-bound and
// slower with it).
static const uint32_t *rec_order(const DevBuf<uint32_t> &ord, long long n) {
  return n >= 16384 ? ord.p : nullptr;
}

// Every element of the array (its leading padding included) set to zero.
template <typename T>
static void zero_buf(DevBuf<T> &b, const int *skip, hipStream_t s) {
  const long long n = (long long)(b.n + b.lead);
  const long long nvec = (long long)(n * sizeof(T) / 16);
  hipLaunchKernelGGL((k_zero<T>), dim3(blocks(std::max(nvec, n - nvec * 16 / (long long)sizeof(T)))),
                     dim3(256), 0, s, b.raw, n, skip);
  WM_LAUNCHED();
}

// getnextcoarselevel3d (f3d_mgrid.F:3601-3652), mgcoarsening = 2.
static void next_coarse(int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                        double dz, int64_t &nxc, int64_t &nyc, int64_t &nzc,
                        double &dxc, double &dyc, double &dzc) {
  nxc = (nx + 1) / 2;
  if (nxc % 2 == 1) nxc = nxc + 1;
  nyc = (ny + 1) / 2;
  if (nyc % 2 == 1) nyc = nyc + 1;
  nzc = (nz + 1) / 2;
  if (nzc % 2 == 1) nzc = nzc + 1;
  double mingridsize = std::min(dx, dy);
  mingridsize = std::min(mingridsize, dz);
  if (dx > 4. / 3. * mingridsize) nxc = nx;
  if (dy > 4. / 3. * mingridsize) nyc = ny;
  if (dz > 4. / 3. * mingridsize) nzc = nz;
  dxc = dx * nx / nxc;
  dyc = nyc == 0 ? dxc : dy * ny / nyc;
  dzc = dz * nz / nzc;
}

static RBGeom make_geom(int nx, int ny, int nz) {
  RBGeom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  // elements per half row: (nx+4)/2 are used; rows of 64 elements or more
  // are padded to a multiple of 16 (128 bytes in FP64) so that, with the
  // arrays' 15-element lead, element 1 of every row starts a cache line
  g.W = (nx + 4) / 2;
  if (g.W >= 64) g.W = (g.W + 15) / 16 * 16;
  g.PY = ny + 3;
  g.PZ = nz + 3;
  g.cs = (long long)g.W * g.PY * g.PZ;
  return g;
}

// restrict3d's per-axis window tables (f3d_mgrid.F:1654-1688), serial
// (offset 0), res guard 1.
static void restrict_axis(int n, int nc, int lbm, int lbp, Level &L, int a) {
  int cmin = lbm == 0 ? 1 : 0, cmax = lbp == 0 ? nc - 1 : nc;
  double di = n > 0 ? 1. * nc / n : 1.;
  std::vector<int> lo(nc + 1, 0), hi(nc + 1, -1);
  std::vector<double> w(5 * (nc + 1), 0.);
  for (int ic = cmin; ic <= cmax; ic++) {
    int imin = ((ic - 1) * n + 4 * nc) / nc - 3;
    int imax = ((ic + 1) * n - 1) / nc;
    lo[ic] = std::max(imin, -1);
    hi[ic] = std::min(imax, n + 1);
    for (int i = imin; i <= imax; i++) {
      WM_REQUIRE(i - imin < 5, "restrict3d window wider than 5 points");
      w[5 * ic + (i - imin)] = 1. - std::fabs(ic - (i + 0.) * di);
    }
  }
  // the plain 2:1 windows [2ic-1, 2ic+1] with weights (1/2, 1, 1/2) over
  // [1, nc-1] (Dirichlet ends)
  bool std21 = n == 2 * nc && lbm == 0 && lbp == 0 && nc >= 2;
  for (int ic = cmin; ic <= cmax && std21; ic++)
    std21 = lo[ic] == 2 * ic - 1 && hi[ic] == 2 * ic + 1 && w[5 * ic] == 0.5 &&
            w[5 * ic + 1] == 1. && w[5 * ic + 2] == 0.5;
  L.rstd[a] = std21;
  L.rlo[a].upload(lo);
  L.rhi[a].upload(hi);
  L.rw[a].upload(w);
  L.rcmin[a] = cmin;
  L.rcmax[a] = cmax;
  L.rwidth[a] = 0;
  for (int ic = cmin; ic <= cmax; ic++) L.rwidth[a] = std::max(L.rwidth[a], hi[ic] - lo[ic] + 1);
  for (int c0 = cmin; c0 <= cmax; c0 += 32)
    L.rspan32[a] = std::max(L.rspan32[a], hi[std::min(c0 + 31, cmax)] - lo[c0] + 1);
  for (int c0 = cmin; c0 <= cmax; c0 += 8)
    L.rspan8[a] = std::max(L.rspan8[a], hi[std::min(c0 + 7, cmax)] - lo[c0] + 1);
  L.rflo[a] = cmax >= cmin ? lo[cmin] : 0;
  L.rfhi[a] = cmax >= cmin ? hi[cmax] : -1;
}

// expand3d's per-axis tables (f3d_mgrid.F:1833-1844).
static void expand_axis(int n, int nc, int lbm, int lbp, Level &L, int a) {
  int lo = lbm == 0 ? 1 : 0, hi = lbp == 0 ? n - 1 : n;
  std::vector<int> j(n + 1, 0);
  std::vector<double> w(n + 1, 0.);
  for (int i = lo; i <= hi; i++) {
    j[i] = (i * nc) / n;
    w[i] = 1. * (i * nc) / n - j[i];
  }
  L.ej[a].upload(j);
  L.ew[a].upload(w);
  L.elo[a] = lo;
  L.ehi[a] = hi;
}

struct LoopRange {
  int ixmin, ixmax, iymin, iymax, izmin, izmax;
};

// Loop limits of sorhalfpass3d (bounds <= 0 excluded, f3d_mgrid.F:2159-2171).
static LoopRange sor_range(const Level &L, const int *lb) {
  LoopRange r{0, L.nx, 0, L.ny, 0, L.nz};
  if (lb[0] <= 0) r.ixmin = 1;
  if (lb[1] <= 0) r.ixmax = L.nx - 1;
  if (lb[2] <= 0) r.iymin = 1;
  if (lb[3] <= 0) r.iymax = L.ny - 1;
  if (lb[4] <= 0) r.izmin = 1;
  if (lb[5] <= 0) r.izmax = L.nz - 1;
  return r;
}

// Loop limits of residual3d and expand3d (bounds == 0 excluded).
static LoopRange res_range(const Level &L, const int *lb) {
  LoopRange r{0, L.nx, 0, L.ny, 0, L.nz};
  if (lb[0] == 0) r.ixmin = 1;
  if (lb[1] == 0) r.ixmax = L.nx - 1;
  if (lb[2] == 0) r.iymin = 1;
  if (lb[3] == 0) r.iymax = L.ny - 1;
  if (lb[4] == 0) r.izmin = 1;
  if (lb[5] == 0) r.izmax = L.nz - 1;
  return r;
}

// Build one colour's subgrid records for the SOR (residual = false) or the
// residual (residual = true) from the reference's condbndymg/condbndyres.
static SubHost build_records(const Level &L, int level, int p, const int *lb,
                             double mgparam, int mgform, int icndbndy,
                             bool residual, bool mgform2init) {
  SubHost h;
  const std::vector<int> &pos = L.subpos_h[p];
  const std::vector<double> &dels = L.subdels_h[p], &volt = L.subvolt_h[p];
  int n = (int)pos.size() / 3;
  double dxsqic, dysqic, dzsqic, dxs, dys, dzs;
  double voltfac;
  if (!residual) {
    double cst = mgparam * 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
    dxsqic = L.dxsqi * cst;
    dysqic = L.dysqi * cst;
    dzsqic = L.dzsqi * cst;
    dxs = dxsqic / mgparam;  // dxsqicos
    dys = dysqic / mgparam;
    dzs = dzsqic / mgparam;
    voltfac = (level == 0 && mgform == 1) ? 1. : 0.;
  } else {
    double cst = 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
    dxs = L.dxsqi * cst;  // dxsqic of condbndyres
    dys = L.dysqi * cst;
    dzs = L.dzsqi * cst;
    voltfac = (level == 0 && (mgform == 1 || mgform2init)) ? 1. : 0.;
  }
  for (int ic = 0; ic < n; ic++) {
    int ix = pos[3 * ic], iy = pos[3 * ic + 1], iz = pos[3 * ic + 2];
    if (ix == 0 && lb[0] < 1) continue;
    if (ix == L.nx && lb[1] < 1) continue;
    if (iy == 0 && lb[2] < 1) continue;
    if (iy == L.ny && lb[3] < 1) continue;
    if (iz == 0 && lb[4] < 1) continue;
    if (iz == L.nz && lb[5] < 1) continue;
    const double *d = &dels[6 * ic], *v = &volt[6 * ic];
    Sym s[6];
    double denom = 1., ppp = 1.;
    bool active;
    if (icndbndy == 1) {
      bool dosub = false;
      const double dsq[3] = {dxs, dys, dzs};
      for (int k = 0; k < 6; k++) {
        if (0 < d[k] && d[k] < 1. && (k < 2 || k > 3 || L.ny > 0)) {
          s[k] = Sym{MODE_CONST, voltfac * v[k] / d[k]};
          denom = denom + (1. - d[k]) / d[k] * dsq[k / 2];
          ppp = std::min(ppp, d[k]);
          dosub = true;
        }
      }
      active = residual ? (ppp < 1.) : dosub;
    } else {
      handlesubgrid2(0, 1, s[0], s[1], dxs, d, v, voltfac, denom, ppp);
      if (L.ny > 0) handlesubgrid2(2, 3, s[2], s[3], dys, d, v, voltfac, denom, ppp);
      handlesubgrid2(4, 5, s[4], s[5], dzs, d, v, voltfac, denom, ppp);
      active = ppp < 1.;
    }
    if (!active) continue;
    uint32_t m = 0, z = 0;
    for (int k = 0; k < 6; k++) m |= (uint32_t)s[k].mode << (2 * k);
    if (residual && icndbndy == 2) {
      for (int k = 0; k < 6; k++)
        if (0. >= d[k] && d[k] >= -1. && (k < 2 || k > 3 || L.ny > 0)) z |= 1u << k;
    }
    h.pos.insert(h.pos.end(), {ix, iy, iz});
    h.modes.push_back(m);
    h.zero.push_back(z);
    for (int k = 0; k < 6; k++) h.k.push_back(s[k].k);
    h.denom.push_back(denom);
    h.fac.push_back(mgform2init ? 1. : ppp);
  }
  // records in red-black element order, so that neighbouring threads of the
  // record kernels touch neighbouring memory
  int nr = h.n();
  std::vector<std::pair<long long, int>> ord(nr);
  for (int i = 0; i < nr; i++)
    ord[i] = {rb_index(L.g, h.pos[3 * i], h.pos[3 * i + 1], h.pos[3 * i + 2]), i};
  std::stable_sort(ord.begin(), ord.end());
  SubHost o;
  for (auto &q : ord) {
    int i = q.second;
    o.pos.insert(o.pos.end(), {h.pos[3 * i], h.pos[3 * i + 1], h.pos[3 * i + 2]});
    o.modes.push_back(h.modes[i]);
    o.zero.push_back(h.zero[i]);
    for (int k = 0; k < 6; k++) o.k.push_back(h.k[6 * i + k]);
    o.denom.push_back(h.denom[i]);
    o.fac.push_back(h.fac[i]);
  }
  return o;
}

template <typename T>
CondDev<T> &cond_of(Level &L);
template <>
CondDev<double> &cond_of<double>(Level &L) { return L.c64; }
template <>
CondDev<float> &cond_of<float>(Level &L) { return L.c32; }

template <typename T>
DevBuf<T> &phi_of(Level &L);
template <>
DevBuf<double> &phi_of<double>(Level &L) { return L.phi64; }
template <>
DevBuf<float> &phi_of<float>(Level &L) { return L.phi32; }
template <>
DevBuf<bf16> &phi_of<bf16>(Level &L) { return L.phi16; }
template <typename T>
DevBuf<T> &phib_of(Level &L);
template <>
DevBuf<double> &phib_of<double>(Level &L) { return L.phi64b; }
template <>
DevBuf<float> &phib_of<float>(Level &L) { return L.phi32b; }
template <>
DevBuf<bf16> &phib_of<bf16>(Level &L) { return L.phi16b; }
template <typename T>
DevBuf<T> &rho_of(Level &L);
template <>
DevBuf<double> &rho_of<double>(Level &L) { return L.rho64; }
template <>
DevBuf<float> &rho_of<float>(Level &L) { return L.rho32; }
template <>
DevBuf<bf16> &rho_of<bf16>(Level &L) { return L.rho16; }

// Storage type of the next coarser level: BF16 is used on the finest level
// only, the levels below it are FP32.
template <typename S>
struct Coarse {
  typedef S type;
};
template <>
struct Coarse<bf16> {
  typedef float type;
};
template <typename T>
DevBuf<T> &res_of(Level &L);
template <>
DevBuf<double> &res_of<double>(Level &L) { return L.res64; }
template <>
DevBuf<float> &res_of<float>(Level &L) { return L.res32; }

struct Params {
  double mgparam;
  int mgform, down, up, lcnd, icnd;
};

// Mask of points the dense kernels leave to the conductor kernels:
// interior points and active subgrid points of this level.
static void build_mask(wm_mg3d *h, Level &L, int level, const Params &P) {
  RecKey key;
  key.mgparam = P.mgparam;  // activity does not depend on it; keep it simple
  key.icndbndy = P.icnd;
  key.lcndbndy = P.lcnd;
  if (L.hasmask && L.maskkey.icndbndy == key.icndbndy && L.maskkey.lcndbndy == key.lcndbndy)
    return;
  std::vector<uint32_t> m((size_t)cdiv(2 * L.g.cs, 32), 0u);
  auto setbit = [&](int ix, int iy, int iz) {
    long long e = rb_index(L.g, ix, iy, iz);
    m[e >> 5] |= 1u << (e & 31);
  };
  std::vector<int> ipos;
  bool any = false;
  if (L.nint > 0) {
    std::vector<int> hp(3 * L.nint);
    WM_HIP(hipMemcpy(hp.data(), L.int_pos.p, sizeof(int) * hp.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < L.nint; i++) setbit(hp[3 * i], hp[3 * i + 1], hp[3 * i + 2]);
    any = true;
    L.imask.upload(m);
  } else {
    L.imask.release();
  }
  L.sub_on_interior = false;
  if (P.lcnd) {
    for (int p = 0; p < 2; p++) {
      SubHost sh = build_records(L, level, p, h->bounds, P.mgparam, 1, P.icnd, false, false);
      for (int i = 0; i < sh.n(); i++) {
        long long e = rb_index(L.g, sh.pos[3 * i], sh.pos[3 * i + 1], sh.pos[3 * i + 2]);
        if ((m[e >> 5] >> (e & 31)) & 1u) L.sub_on_interior = true;
        setbit(sh.pos[3 * i], sh.pos[3 * i + 1], sh.pos[3 * i + 2]);
        any = true;
      }
    }
  }
  if (any) L.mask.upload(m);
  else L.mask.release();
  L.hasmask = any;
  L.maskkey = key;
}

template <typename T>
static void ensure_records(wm_mg3d *h, Level &L, int level, const Params &P,
                           bool residual, int mgform2init) {
  CondDev<T> &C = cond_of<T>(L);
  RecKey key;
  key.mgparam = P.mgparam;
  // voltages enter only on the finest level (voltfac), so the records of
  // coarser levels are the same for both mgforms and are built once
  key.mgform = level == 0 ? P.mgform : 0;
  key.icndbndy = P.icnd;
  key.lcndbndy = P.lcnd;
  key.mgform2init = residual ? mgform2init : 0;
  if (!residual) {
    if (C.sorkey == key) return;
    for (int p = 0; p < 2; p++) {
      SubHost sh;
      if (P.lcnd) sh = build_records(L, level, p, h->bounds, P.mgparam, P.mgform, P.icnd, false, false);
      C.sor[p].upload(sh);
    }
    C.sorkey = key;
  } else {
    if (C.reskey[mgform2init] == key) return;
    std::vector<std::pair<long long, uint32_t>> ord;
    for (int p = 0; p < 2; p++) {
      SubHost sh;
      if (P.lcnd)
        sh = build_records(L, level, p, h->bounds, P.mgparam, P.mgform, P.icnd, true, mgform2init != 0);
      for (int i = 0; i < sh.n(); i++) {
        const long long nat = (long long)sh.pos[3 * i] + 1 +
                              (long long)(L.nx + 3) * ((long long)sh.pos[3 * i + 1] + 1 +
                                                       (long long)(L.ny + 3) * (sh.pos[3 * i + 2] + 1));
        ord.push_back({nat, ((uint32_t)p << 31) | (uint32_t)i});
      }
      C.res[mgform2init][p].upload(sh);
    }
    std::sort(ord.begin(), ord.end());
    std::vector<uint32_t> o(ord.size());
    for (size_t i = 0; i < ord.size(); i++) o[i] = ord[i].second;
    C.resord[mgform2init].upload(o);
    C.reskey[mgform2init] = key;
  }
}

template <typename T>
static void ensure_buffers(Level &L, bool with_res) {
  phi_of<T>(L).alloc_rb(2 * L.g.cs);
  phib_of<T>(L).alloc_rb(2 * L.g.cs);
  rho_of<T>(L).alloc_rb(2 * L.g.cs);
  if (with_res) res_of<T>(L).alloc_rb(std::max<size_t>(2 * L.g.cs, (size_t)(L.nx + 3) * (L.ny + 3) * (L.nz + 3)));
}

template <typename T, typename Idx>
static void launch_bc(T *u, Idx I, int nx, int ny, int nz, const int *b, int lwd, int lzd,
                      const int *skip, hipStream_t s) {
  // x, then y, then z; skipped entirely when nothing would change.
  int need[3];
  for (int a = 0; a < 3; a++)
    need[a] = (b[2 * a] != 0 || b[2 * a + 1] != 0 || lwd || lzd);
  int dims[3] = {nx, ny, nz};
  for (int a = 0; a < 3; a++) {
    if (!need[a]) continue;
    int na = a == 0 ? ny : nx, nb = a == 2 ? ny : nz;
    long long n = (long long)(na + 3) * (nb + 3);
    hipLaunchKernelGGL((k_bc_axis<T, Idx>), dim3(blocks(n)), dim3(256), 0, s, u, I, a, nx, ny,
                       nz, b[2 * a], b[2 * a + 1], lwd, lzd, skip);
    WM_LAUNCHED();
    (void)dims;
  }
}

template <typename T>
static void pin_and_mirror(wm_mg3d *h, Level &L, int level, const Params &P, T *phi,
                           int usevolt, const int *skip, hipStream_t s) {
  if (L.nint > 0) {
    hipLaunchKernelGGL((k_pin<T>), dim3(blocks(L.nint)), dim3(256), 0, s, phi, L.g, L.nint,
                       L.int_pos.p, L.int_volt.p, usevolt, skip);
    WM_LAUNCHED();
  }
  if (P.icnd == 2 && L.neumann) {
    const int *b = h->bounds;
    hipLaunchKernelGGL((k_condbndymgint<T>), dim3(1), dim3(64), 0, s, phi, L.g,
                       (int)(L.subpos_h[0].size() / 3), (int)(L.subpos_h[1].size() / 3),
                       L.subpos_d[0].p, L.subpos_d[1].p, L.subdels_d[0].p, L.subdels_d[1].p,
                       b[0], b[1], b[2], b[3], b[4], b[5], skip);
    WM_LAUNCHED();
  }
}

// sorhalfpass3d (f3d_mgrid.F:2095-2291) on level data in red-black layout.
// src (not NULL): read the values before the sweep from src, write phi
// (k_sor_zmarch); wsave false: the convergence test leaves save unchanged.
template <typename S>
static void halfpass(wm_mg3d *h, Level &L, int level, const Params &P, int c,
                     const int *skip, hipStream_t s, bool inpass = true,
                     double *cpart = nullptr, double *cpart_end = nullptr,
                     int *nslots = nullptr, const S *src = nullptr, bool wsave = true) {
  typedef typename St<S>::T T;
  S *phi = phi_of<S>(L).p;
  const S *rho = rho_of<S>(L).p;
  LoopRange r = sor_range(L, h->bounds);
  double cst = P.mgparam * 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
  T dxsqic = (T)(L.dxsqi * cst), dysqic = (T)(L.dysqi * cst), dzsqic = (T)(L.dzsqi * cst);
  T spm1 = (T)(1. - P.mgparam);
  SubList<T> sub{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (P.lcnd) sub = cond_of<T>(L).sor[c].view();
  int nyr = r.iymax - r.iymin + 1, nzr = r.izmax - r.izmin + 1;
  if (nyr <= 0 || nzr <= 0) return;
  // rows per z-marching block (wavefronts of 64 lanes, one row each)
  static const int rows_env = getenv("WARPMI_SOR_ROWS") ? atoi(getenv("WARPMI_SOR_ROWS")) : 0;
  static const bool tiled_env0 = getenv("WARPMI_SOR_TILED") != nullptr;
  int rows = rows_env > 0 && rows_env <= 16 && !tiled_env0 ? rows_env : 4;
  int nthr = 64 * rows;
  int nby = (int)cdiv(nyr, rows);
  int nbsub = (int)(8 * cdiv(cdiv(sub.n, nthr), 8));
  static const bool tiled_env = getenv("WARPMI_SOR_TILED") != nullptr;
  static const long long tiled_below =
      getenv("WARPMI_TILED_BELOW") ? atoll(getenv("WARPMI_TILED_BELOW")) : 0;
  const long long nodes = (long long)(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
  const bool tiled = (tiled_env || (cpart == nullptr && nodes < tiled_below)) &&
                     !std::is_same<S, bf16>::value && src == nullptr;
  // z-marching sweep: colour elements j0..j1 cover ix in [ixmin, ixmax]
  // for both row parities; z chunks sized for about 2048 dense blocks.
  int j0 = (r.ixmin + 1) / 2, j1 = (r.ixmax + 1) / 2;
  int nbx = tiled ? (int)cdiv(L.g.W, 64) : (int)cdiv(j1 - j0 + 1, 64);
  long long cols = (long long)nbx * nby;
  static const int zc_env = getenv("WARPMI_SOR_ZC") ? atoi(getenv("WARPMI_SOR_ZC")) : 0;
  int zc = zc_env > 0 ? zc_env
                      : (int)std::min<long long>(64, std::max<long long>(2, cdiv((long long)nzr * cols, 2048)));
  int nbz = (int)cdiv(nzr, zc);
  int nbdense = tiled ? nbx * nby * nzr : nbx * nby * nbz;
  // the event timing covers the plain finest-level sweeps (the sweeps that
  // carry the convergence test move more bytes per point)
  bool timed = h->prof.on && level == 0 && cpart == nullptr;
  if (timed) {
    while ((int)h->prof.ev.size() < 2 * (h->prof.used + 1)) {
      hipEvent_t e;
      WM_HIP(hipEventCreate(&e));
      h->prof.ev.push_back(e);
    }
    WM_HIP(hipEventRecord(h->prof.ev[2 * h->prof.used], s));
  }
  if constexpr (!std::is_same<S, bf16>::value) {
    if (tiled)
      hipLaunchKernelGGL((k_sor_halfpass<T>), dim3(nbdense + (int)cdiv(sub.n, 256)), dim3(256), 0,
                         s, phi, rho,
                         L.g, c, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin, r.izmax, dxsqic,
                         dysqic, dzsqic, spm1, L.hasmask ? L.mask.p : nullptr, sub, nbx, nby,
                         nbdense, skip);
  }
  if (cpart) {
    // the sweep with the convergence test folded in (FP64 only)
    long long ns = (long long)(nbdense + nbsub) * (nthr / 64);
    const bool fits = !tiled && std::is_same<S, double>::value && ns <= cpart_end - cpart;
    WM_REQUIRE(fits, "sorhalfpass3d: the fused convergence test does not fit this sweep");
    // the dense sweep reads mask words only when the level has a mask
    if constexpr (std::is_same<S, double>::value) {
      if (L.hasmask)
        hipLaunchKernelGGL((k_sor_zmarch<S, T, true, true>), dim3(nbdense + nbsub), dim3(nthr),
                           0, s, phi, rho, L.g, c, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin,
                           r.izmax, dxsqic, dysqic, dzsqic, spm1, (const uint32_t *)L.mask.p, sub,
                           j0, nbx, nby, zc, nbdense, nbsub, h->save64.p, cpart, skip, src,
                           (int)wsave);
      else
        hipLaunchKernelGGL((k_sor_zmarch<S, T, true, false>), dim3(nbdense + nbsub), dim3(nthr),
                           0, s, phi, rho, L.g, c, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin,
                           r.izmax, dxsqic, dysqic, dzsqic, spm1, (const uint32_t *)nullptr, sub,
                           j0, nbx, nby, zc, nbdense, nbsub, h->save64.p, cpart, skip, src,
                           (int)wsave);
    }
    *nslots = (int)ns;
  } else if (!tiled) {
    if (L.hasmask)
      hipLaunchKernelGGL((k_sor_zmarch<S, T, false, true>), dim3(nbdense + nbsub), dim3(nthr), 0,
                         s, phi, rho, L.g, c, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin,
                         r.izmax, dxsqic, dysqic, dzsqic, spm1, (const uint32_t *)L.mask.p, sub,
                         j0, nbx, nby, zc, nbdense, nbsub, (S *)nullptr, (double *)nullptr, skip,
                         src, 1);
    else
      hipLaunchKernelGGL((k_sor_zmarch<S, T, false, false>), dim3(nbdense + nbsub), dim3(nthr),
                         0, s, phi, rho, L.g, c, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin,
                         r.izmax, dxsqic, dysqic, dzsqic, spm1, (const uint32_t *)nullptr, sub,
                         j0, nbx, nby, zc, nbdense, nbsub, (S *)nullptr, (double *)nullptr, skip,
                         src, 1);
  }
  WM_LAUNCHED();
  if (timed) {
    WM_HIP(hipEventRecord(h->prof.ev[2 * h->prof.used + 1], s));
    h->prof.used++;
  }
  // cond_potmg after the sweep (f3d_mgrid.F:2276-2285).  The dense sweep
  // skips every interior point and the pass starts with a pin, so the
  // pin is a no-op unless a subgrid record or a periodic wrap can write an
  // interior point, or Neumann subgrid ghosts must be mirrored.  A half
  // sweep called alone (inpass false) always pins.
  const int *b = h->bounds;
  bool periodic = b[0] == 2 || b[2] == 2 || b[4] == 2;
  if (!inpass || L.sub_on_interior || periodic || (P.icnd == 2 && L.neumann)) {
    int usevolt = (level == 0 && P.mgform == 1);
    pin_and_mirror<S>(h, L, level, P, phi, usevolt, skip, s);
  }
}

// Shape of the fused pass on a level: lanes per row (one per colour element
// of the row's interior, in whole wavefronts), owned rows per block, the z
// chunk and the number of blocks.
struct FusedShape {
  int j0, nj, nl, R, zc, nby, nbz;
};

// Resident blocks per compute unit of a k_sor_fused form for a block size
// and LDS size, from the runtime's occupancy calculator (kept for the last
// shape asked).
template <typename S, bool CONV, bool MASK>
static int fused_occupancy(int nthr, size_t lds) {
  static int n = -1, nthr_of_n = 0;
  static size_t lds_of_n = 0;
  if (n < 0 || lds != lds_of_n || nthr != nthr_of_n) {
    int k = 0;
    WM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &k, reinterpret_cast<const void *>(&k_sor_fused<S, CONV, MASK>), nthr, lds));
    n = std::max(k, 1);
    lds_of_n = lds;
    nthr_of_n = nthr;
  }
  return n;
}

static FusedShape fused_shape(const wm_mg3d *h, const Level &L, int per_cu) {
  LoopRange r = sor_range(L, h->bounds);
  FusedShape f{};
  f.j0 = (r.ixmin + 1) / 2;
  f.nj = (r.ixmax + 1) / 2 - f.j0 + 1;
  f.nl = (int)(64 * cdiv(std::max(f.nj, 1), 64));
  static const int rows_env = getenv("WARPMI_FUSED_ROWS") ? atoi(getenv("WARPMI_FUSED_ROWS")) : 0;
  int rows = 1024 / f.nl;  // rows per block, halo rows included
  if (rows_env >= 3 && rows_env < rows) rows = rows_env;
  f.R = std::max(1, rows - 2);
  int nyr = r.iymax - r.iymin + 1, nzr = std::max(r.izmax - r.izmin + 1, 1);
  f.nby = (int)cdiv(std::max(nyr, 1), f.R);
  // one round of resident blocks: as many z chunks as the row tiles leave
  // room for (at least 4 planes each)
  static const int zc_env = getenv("WARPMI_FUSED_ZC") ? atoi(getenv("WARPMI_FUSED_ZC")) : 0;
  long long slots = (long long)per_cu * h->ncu;
  long long nbz = std::max<long long>(1, slots / f.nby);
  int zc = (int)std::max<long long>(4, cdiv(nzr, nbz));
  f.zc = zc_env > 0 ? zc_env : std::min(zc, nzr);
  f.nbz = (int)cdiv(nzr, f.zc);
  return f;
}

// Whether a pass on this level can be the fused two-colour pass: six
// Dirichlet boundaries, no record on an interior point, no Neumann subgrid
// mirror, rows of at most 256 colour elements (two rows of halo and at
// least one owned row per 1024-thread block), an even number of passes per
// V-cycle (so that phi ends each V-cycle in its first buffer).
static bool fused_ok(const wm_mg3d *h, const Level &L, int level, const Params &P) {
  if (h->sor_form == 0 || (h->sor_form == 2 && level != 0) || getenv("WARPMI_SOR_TILED"))
    return false;
  for (int i = 0; i < 6; i++)
    if (h->bounds[i] != 0) return false;
  if (L.sub_on_interior || (P.icnd == 2 && L.neumann)) return false;
  if ((P.down + P.up) % 2 != 0) return false;
  LoopRange r = sor_range(L, h->bounds);
  int nj = (r.ixmax + 1) / 2 - (r.ixmin + 1) / 2 + 1;
  return nj >= 1 && nj <= 256 && r.iymax >= r.iymin && r.izmax >= r.izmin;
}

// The subgrid records of colour c alone (a record-only k_sor_zmarch) on
// the red-black array phi; with cpart, the convergence test as well.
// Returns the number of partial-maximum slots written.
template <typename S>
static int records_only(Level &L, const Params &P, int c, S *phi, const LoopRange &r,
                        typename St<S>::T dxsqic, typename St<S>::T dysqic,
                        typename St<S>::T dzsqic, typename St<S>::T spm1, const int *skip,
                        hipStream_t s, S *save, double *cpart) {
  typedef typename St<S>::T T;
  if (!P.lcnd) return 0;
  SubList<T> sub = cond_of<T>(L).sor[c].view();
  if (sub.n == 0) return 0;
  const int nthr = 256;
  int nbsub = (int)cdiv(sub.n, nthr);
  const uint32_t *mask = L.hasmask ? L.mask.p : nullptr;
  if (cpart) {
    if constexpr (std::is_same<S, double>::value)
      hipLaunchKernelGGL((k_sor_zmarch<S, T, true>), dim3(nbsub), dim3(nthr), 0, s, phi,
                         (const S *)rho_of<S>(L).p, L.g, c, r.ixmin, r.ixmax, r.iymin, r.iymax,
                         r.izmin, r.izmax, dxsqic, dysqic, dzsqic, spm1, mask, sub, 0, 1, 1, 1,
                         0, nbsub, save, cpart, skip, (const S *)nullptr, 1);
  } else {
    hipLaunchKernelGGL((k_sor_zmarch<S>), dim3(nbsub), dim3(nthr), 0, s, phi,
                       (const S *)rho_of<S>(L).p, L.g, c, r.ixmin, r.ixmax, r.iymin, r.iymax,
                       r.izmin, r.izmax, dxsqic, dysqic, dzsqic, spm1, mask, sub, 0, 1, 1, 1, 0,
                       nbsub, (S *)nullptr, (double *)nullptr, skip, (const S *)nullptr, 1);
  }
  WM_LAUNCHED();
  return cpart ? nbsub * (nthr / 64) : 0;
}

// The fused pass: colour-0 records in place in the first buffer, the dense
// pass from the first buffer into the second, colour-1 records in the
// second, then the buffers exchanged.  cpart as in halfpass (FP64 only).
template <typename S>
static void fused_pass(wm_mg3d *h, Level &L, int level, const Params &P, const int *skip,
                       hipStream_t s, double *cpart, double *cpart_end, int *nslots) {
  typedef typename St<S>::T T;
  DevBuf<S> &A = phi_of<S>(L), &B = phib_of<S>(L);
  LoopRange r = sor_range(L, h->bounds);
  double cst = P.mgparam * 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
  T dxsqic = (T)(L.dxsqi * cst), dysqic = (T)(L.dysqi * cst), dzsqic = (T)(L.dzsqi * cst);
  T spm1 = (T)(1. - P.mgparam);
  const bool conv = cpart != nullptr;
  const uint32_t *mask = L.hasmask ? L.mask.p : nullptr;
  FusedShape f0 = fused_shape(h, L, 1);
  size_t lds = sizeof(S) * 3 * (size_t)(f0.R + 2) * (size_t)(f0.nj + 2);
  // a thread per lane of every row of the block, halo rows included
  const int nthr = (f0.R + 2) * f0.nl;
  int per_cu = 1;
  if (conv) {
    if constexpr (std::is_same<S, double>::value)
      per_cu = mask ? fused_occupancy<S, true, true>(nthr, lds)
                    : fused_occupancy<S, true, false>(nthr, lds);
  } else {
    per_cu = mask ? fused_occupancy<S, false, true>(nthr, lds)
                  : fused_occupancy<S, false, false>(nthr, lds);
  }
  FusedShape f = fused_shape(h, L, per_cu);
  WM_REQUIRE(f.R == f0.R && (f.R + 2) * f.nl == nthr && nthr <= 1024,
             "sorpass3d: inconsistent fused-pass tile");
  int nb = f.nby * f.nbz;
  if (conv) {
    const bool fits = std::is_same<S, double>::value &&
                      (long long)nb * 16 + 8 * (cdiv(cond_of<T>(L).sor[0].n, 256) +
                                                cdiv(cond_of<T>(L).sor[1].n, 256)) <=
                          cpart_end - cpart;
    WM_REQUIRE(fits, "sorpass3d: the fused convergence test does not fit this pass");
  }
  bool timed = h->prof.on && level == 0 && cpart == nullptr;
  if (timed) {
    while ((int)h->prof.ev.size() < 2 * (h->prof.used + 1)) {
      hipEvent_t e;
      WM_HIP(hipEventCreate(&e));
      h->prof.ev.push_back(e);
    }
    WM_HIP(hipEventRecord(h->prof.ev[2 * h->prof.used], s));
    h->prof.fused = true;
  }
  int used = records_only<S>(L, P, 0, A.p, r, dxsqic, dysqic, dzsqic, spm1, skip, s,
                             conv ? (S *)h->save64.p : nullptr, cpart);
#define WM_FUSED_LAUNCH(CV, MK)                                                                \
  hipLaunchKernelGGL((k_sor_fused<S, CV, MK>), dim3(nb), dim3(nthr), lds, s, (const S *)A.p, \
                     B.p, (const S *)rho_of<S>(L).p, L.g, r.ixmin, r.ixmax, r.iymin, r.iymax, \
                     r.izmin, r.izmax, dxsqic, dysqic, dzsqic, spm1, mask, f.j0, f.nj, f.nl,   \
                     f.R, f.zc, f.nby, nb, CV ? (S *)h->save64.p : (S *)nullptr,              \
                     CV ? cpart + used : (double *)nullptr, skip)
  if (conv) {
    if constexpr (std::is_same<S, double>::value) {
      if (mask)
        WM_FUSED_LAUNCH(true, true);
      else
        WM_FUSED_LAUNCH(true, false);
    }
    used += nb * (nthr / 64);
  } else if (mask) {
    WM_FUSED_LAUNCH(false, true);
  } else {
    WM_FUSED_LAUNCH(false, false);
  }
#undef WM_FUSED_LAUNCH
  WM_LAUNCHED();
  used += records_only<S>(L, P, 1, B.p, r, dxsqic, dysqic, dzsqic, spm1, skip, s,
                          conv ? (S *)h->save64.p : nullptr, conv ? cpart + used : nullptr);
  if (timed) {
    WM_HIP(hipEventRecord(h->prof.ev[2 * h->prof.used + 1], s));
    h->prof.used++;
  }
  if (nslots) *nslots = used;
  A.swap(B);
}

// src (not NULL, half sweeps only): the pass starts from the values in src
// and writes phi; wsave false: the convergence test leaves save unchanged.
template <typename T>
static void sorpass(wm_mg3d *h, Level &L, int level, const Params &P, const int *skip,
                    hipStream_t s, bool pin = true, double *cpart = nullptr,
                    double *cpart_end = nullptr, int *nslots = nullptr,
                    const T *src = nullptr, bool wsave = true) {
  int usevolt = (level == 0 && P.mgform == 1);
  if (fused_ok(h, L, level, P) && phib_of<T>(L).p) {
    WM_REQUIRE(!src && wsave, "sorpass: the fused pass has no second-buffer form");
    // the pin writes both buffers, so that conductor points on the mesh
    // boundary (which no pass writes) agree in both
    if (pin) {
      pin_and_mirror<T>(h, L, level, P, phi_of<T>(L).p, usevolt, skip, s);
      pin_and_mirror<T>(h, L, level, P, phib_of<T>(L).p, usevolt, skip, s);
    }
    int ns = 0;
    fused_pass<T>(h, L, level, P, skip, s, cpart, cpart_end, &ns);
    if (nslots) *nslots = ns;
    return;
  }
  WM_REQUIRE(!src || !pin, "sorpass: a pass from a second buffer cannot start with a pin");
  if (pin) pin_and_mirror<T>(h, L, level, P, phi_of<T>(L).p, usevolt, skip, s);
  int used = 0;
  for (int c = 0; c <= 1; c++) {
    int ns = 0;
    halfpass<T>(h, L, level, P, c, skip, s, true, cpart ? cpart + used : nullptr, cpart_end,
                &ns, src, wsave);
    used += ns;
    launch_bc<T>(phi_of<T>(L).p, RBIdx{L.g}, L.nx, L.ny, L.nz, h->bounds, 0, 0, skip, s);
  }
  if (nslots) *nslots = used;
}

// residual3d (f3d_mgrid.F:3367-3598) from red-black phi/rho into the
// red-black array res (storage TR, same geometry): the points outside the
// loop range are zeroed (k_res_shell), the dense points computed in one
// z-marching launch for both colours, then cond_potmgres/condbndyres
// (k_residual_cond, even list then odd list) and the boundary conditions.
//
// steady: res already holds this solve's zeros outside the loop range and
// at the interior conductor points (residual_zeros), which nothing but
// the boundary conditions below writes between calls; then only the dense
// points and the records are computed, the records of both colours in one
// launch unless one of them zeroes a neighbour's residual.
template <typename S, typename TR>
static void residual_rb(wm_mg3d *h, Level &L, int level, const Params &P, const S *phi,
                        const S *rho, TR *res, int mgform2init, const int *skip, hipStream_t s,
                        bool steady = false) {
  typedef typename St<S>::T T;
  LoopRange r = res_range(L, h->bounds);
  double cst = 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
  T dxsqic = (T)(L.dxsqi * P.mgparam * cst), dysqic = (T)(L.dysqi * P.mgparam * cst),
    dzsqic = (T)(L.dzsqi * P.mgparam * cst);
  long long nsh = 4ll * ((long long)(L.ny + 3) * (L.nz + 3) + (long long)(L.nx + 3) * (L.nz + 3) +
                         (long long)(L.nx + 3) * (L.ny + 3));
  if (!steady) {
    hipLaunchKernelGGL((k_res_shell<TR>), dim3(blocks(nsh)), dim3(256), 0, s, res, L.g, r.ixmin,
                       r.ixmax, r.iymin, r.iymax, r.izmin, r.izmax, skip);
    WM_LAUNCHED();
  }
  int nyr = r.iymax - r.iymin + 1, nzr = r.izmax - r.izmin + 1;
  if (nyr > 0 && nzr > 0 && r.ixmax >= r.ixmin) {
    int nby = (int)cdiv(nyr, 4);
    int j0 = (r.ixmin + 1) / 2, j1 = (r.ixmax + 1) / 2;
    int nbx = (int)cdiv(j1 - j0 + 1, 64);
    long long cols = (long long)nbx * nby;
    int zc = (int)std::min<long long>(64, std::max<long long>(2, cdiv((long long)nzr * cols, 2048)));
    int nbz = (int)cdiv(nzr, zc);
    int nbdense = nbx * nby * nbz;
    WM_REQUIRE(2 * L.g.cs < (1ll << 31), "residual3d: level too large for 32-bit indexing");
    static const bool one_env = getenv("WARPMI_RESIDUAL_ONE_PAIR") != nullptr;
    bool launched = false;
    if constexpr (std::is_same<S, double>::value) {
      if (!one_env && L.g.W >= 64 && L.g.W % 16 == 0) {
        // two pairs per thread from an odd element (16-byte aligned)
        const int jb = (j0 & 1) ? j0 : j0 - 1;
        const int nbx2 = (int)cdiv(j1 - jb + 1, 128);
        const long long cols2 = (long long)nbx2 * nby;
        const int zc2 = (int)std::min<long long>(
            64, std::max<long long>(2, cdiv((long long)nzr * cols2, 2048)));
        const int nb2 = nbx2 * nby * (int)cdiv(nzr, zc2);
        if (L.hasmask)
          hipLaunchKernelGGL((k_residual_pairs2<TR, true>), dim3(nb2), dim3(256), 0, s, phi, rho,
                             res, L.g, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin, r.izmax,
                             (double)dxsqic, (double)dysqic, (double)dzsqic, P.mgparam,
                             (const uint32_t *)L.mask.p, jb, j1, nbx2, nby, zc2, nb2, skip);
        else
          hipLaunchKernelGGL((k_residual_pairs2<TR, false>), dim3(nb2), dim3(256), 0, s, phi, rho,
                             res, L.g, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin, r.izmax,
                             (double)dxsqic, (double)dysqic, (double)dzsqic, P.mgparam,
                             (const uint32_t *)nullptr, jb, j1, nbx2, nby, zc2, nb2, skip);
        WM_LAUNCHED();
        launched = true;
      }
    }
    if (!launched) {
      if (L.hasmask)
        hipLaunchKernelGGL((k_residual_pairs<S, TR, T, true>), dim3(nbdense), dim3(256), 0, s,
                           phi, rho, res, L.g, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin,
                           r.izmax, dxsqic, dysqic, dzsqic, (T)P.mgparam,
                           (const uint32_t *)L.mask.p, j0, j1, nbx, nby, zc, nbdense, skip);
      else
        hipLaunchKernelGGL((k_residual_pairs<S, TR, T, false>), dim3(nbdense), dim3(256), 0, s,
                           phi, rho, res, L.g, r.ixmin, r.ixmax, r.iymin, r.iymax, r.izmin,
                           r.izmax, dxsqic, dysqic, dzsqic, (T)P.mgparam,
                           (const uint32_t *)nullptr, j0, j1, nbx, nby, zc, nbdense, skip);
      WM_LAUNCHED();
    }
  }
  if (L.hasmask) {
    ensure_records<T>(h, L, level, P, true, mgform2init);
    CondDev<T> &C = cond_of<T>(L);
    T dxsqics = (T)(P.mgparam * L.dxsqi * cst), dysqics = (T)(P.mgparam * L.dysqi * cst),
      dzsqics = (T)(P.mgparam * L.dzsqi * cst);
    if (steady) {
      SubList<T> s0 = C.res[mgform2init][0].view(), s1 = C.res[mgform2init][1].view();
      if (!C.res[mgform2init][0].anyzero && !C.res[mgform2init][1].anyzero) {
        if (s0.n + s1.n > 0) {
          hipLaunchKernelGGL((k_residual_rec2<S, TR, T, RBIdx>), dim3(blocks(s0.n + s1.n)),
                             dim3(256), 0, s, phi, rho, res, RBIdx{L.g}, L.g, s0, s1, dxsqics,
                             dysqics, dzsqics, (T)P.mgparam, skip,
                             rec_order(C.resord[mgform2init], s0.n + s1.n));
          WM_LAUNCHED();
        }
      } else {
        const int *none = nullptr;
        for (int p = 0; p < 2; p++) {
          SubList<T> sub = p ? s1 : s0;
          if (sub.n == 0) continue;
          hipLaunchKernelGGL((k_residual_cond<S, TR, T, RBIdx>), dim3(blocks(sub.n)), dim3(256), 0,
                             s, phi, rho, res, RBIdx{L.g}, L.g, 0, none, sub, dxsqics, dysqics,
                             dzsqics, (T)P.mgparam, skip);
          WM_LAUNCHED();
        }
      }
      launch_bc<TR>(res, RBIdx{L.g}, L.nx, L.ny, L.nz, h->bounds, 0, 0, skip, s);
      return;
    }
    // interior points get 0 before any record is evaluated (cond_potmgres
    // then condbndyres): in one launch unless a record sits on an interior
    // point
    bool split = L.sub_on_interior && L.nint > 0;
    if (split) {
      SubList<T> none{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
      hipLaunchKernelGGL((k_residual_cond<S, TR, T, RBIdx>), dim3(blocks(L.nint)), dim3(256), 0,
                         s, phi, rho, res, RBIdx{L.g}, L.g, L.nint, L.int_pos.p, none, dxsqics,
                         dysqics, dzsqics, (T)P.mgparam, skip);
      WM_LAUNCHED();
    }
    for (int p = 0; p < 2; p++) {
      SubList<T> sub = C.res[mgform2init][p].view();
      int nint = p == 0 && !split ? L.nint : 0;
      long long cnt = (long long)nint + sub.n;
      if (cnt == 0) continue;
      hipLaunchKernelGGL((k_residual_cond<S, TR, T, RBIdx>), dim3(blocks(cnt)), dim3(256), 0, s,
                         phi, rho, res, RBIdx{L.g}, L.g, nint, L.int_pos.p, sub, dxsqics,
                         dysqics, dzsqics, (T)P.mgparam, skip);
      WM_LAUNCHED();
    }
  }
  launch_bc<TR>(res, RBIdx{L.g}, L.nx, L.ny, L.nz, h->bounds, 0, 0, skip, s);
}

// The points of res that residual_rb's steady form leaves alone: zeros
// outside the loop range (k_res_shell) and at the interior conductor points
// (cond_potmgres).  Once per solve, before the iterations.
template <typename TR>
static void residual_zeros(wm_mg3d *h, Level &L, TR *res, hipStream_t s) {
  LoopRange r = res_range(L, h->bounds);
  long long nsh = 4ll * ((long long)(L.ny + 3) * (L.nz + 3) + (long long)(L.nx + 3) * (L.nz + 3) +
                         (long long)(L.nx + 3) * (L.ny + 3));
  hipLaunchKernelGGL((k_res_shell<TR>), dim3(blocks(nsh)), dim3(256), 0, s, res, L.g, r.ixmin,
                     r.ixmax, r.iymin, r.iymax, r.izmin, r.izmax, (const int *)nullptr);
  WM_LAUNCHED();
  if (L.hasmask && L.nint > 0) {
    SubList<double> none{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL((k_residual_cond<double, TR, double, RBIdx>), dim3(blocks(L.nint)),
                       dim3(256), 0, s, (const double *)nullptr, (const double *)nullptr, res,
                       RBIdx{L.g}, L.g, L.nint, L.int_pos.p, none, 0., 0., 0., 0.,
                       (const int *)nullptr);
    WM_LAUNCHED();
  }
}

struct RhoIdx {
  int nx, ny;
  __device__ long long operator()(int ix, int iy, int iz) const {
    return (long long)ix + (long long)(nx + 1) * ((long long)iy + (long long)(ny + 1) * iz);
  }
};

template <typename TR, typename TC, typename OutIdx, typename FineIdx>
static void restrict_to(Level &L, const TR *res, FineIdx F, TC *out, OutIdx O, const int *skip,
                        hipStream_t s, int nxc, int nyc, int nzc) {
  RestrictAxis ax{L.rlo[0].p, L.rhi[0].p, L.rw[0].p, L.rcmin[0], L.rcmax[0]};
  RestrictAxis ay{L.rlo[1].p, L.rhi[1].p, L.rw[1].p, L.rcmin[1], L.rcmax[1]};
  RestrictAxis az{L.rlo[2].p, L.rhi[2].p, L.rw[2].p, L.rcmin[2], L.rcmax[2]};
  long long n = (long long)(nxc + 1) * (nyc + 1) * (nzc + 1);
  if constexpr (std::is_same<TC, double>::value) {
    hipLaunchKernelGGL((k_restrict<TR, TC, OutIdx, FineIdx>), dim3(blocks(n)), dim3(256), 0, s,
                       res, F, out, O, nxc, nyc, nzc, ax, ay, az, L.ff, skip);
  } else {
    hipLaunchKernelGGL((k_restrict_t<TC, OutIdx, FineIdx>), dim3(blocks(n)), dim3(256), 0, s,
                       res, F, out, O, nxc, nyc, nzc, ax, ay, az, (TC)L.ff, skip);
  }
  WM_LAUNCHED();
}

template <typename T, typename FI, typename CI>
static void expand_from(Level &L, T *phi, FI F, const T *pc, CI Cc, const int *skip,
                        hipStream_t s) {
  ExpandAxis ax{L.ej[0].p, L.ew[0].p, L.elo[0], L.ehi[0]};
  ExpandAxis ay{L.ej[1].p, L.ew[1].p, L.elo[1], L.ehi[1]};
  ExpandAxis az{L.ej[2].p, L.ew[2].p, L.elo[2], L.ehi[2]};
  long long n = (long long)std::max(0, ax.hi - ax.lo + 1) * std::max(0, ay.hi - ay.lo + 1) *
                std::max(0, az.hi - az.lo + 1);
  if (n == 0) return;
  hipLaunchKernelGGL((k_expand<T, FI, CI>), dim3(blocks(n)), dim3(256), 0, s, phi, F, pc, Cc,
                     ax, ay, az, skip);
  WM_LAUNCHED();
}


// On a level without voltages (an error level: coarser than the finest, or
// the finest in the residual-correction form) the pin that starts the
// first pass after the expansion writes zero at the interior conductor
// points.  When nothing comes between the two -- six Dirichlet boundaries
// (launch_bc does nothing) and no Neumann subgrid mirror -- the expansion
// writes that zero itself and the pin is left out: the values are those of
// expand-then-pin.
static bool pin_fusable(const wm_mg3d *h, const Level &L, int level, const Params &P) {
  const int *b = h->bounds;
  for (int i = 0; i < 6; i++)
    if (b[i] != 0) return false;
  const bool usevolt = (level == 0 && P.mgform == 1);
  return !usevolt && L.nint > 0 && L.imask.p != nullptr && !(P.icnd == 2 && L.neumann);
}

template <typename S, typename SC = typename Coarse<S>::type>
static void expand_rb(wm_mg3d *h, Level &L, Level &Lc, int level, const Params &P,
                      const int *skip, hipStream_t s, bool fusepin = false) {
  ExpandAxis ax{L.ej[0].p, L.ew[0].p, L.elo[0], L.ehi[0]};
  ExpandAxis ay{L.ej[1].p, L.ew[1].p, L.elo[1], L.ehi[1]};
  ExpandAxis az{L.ej[2].p, L.ew[2].p, L.elo[2], L.ehi[2]};
  long long rows = 2ll * std::max(0, ay.hi - ay.lo + 1) * std::max(0, az.hi - az.lo + 1);
  if (rows == 0 || ax.hi < ax.lo) {
    if (fusepin) pin_and_mirror<S>(h, L, level, P, phi_of<S>(L).p, 0, skip, s);
    return;
  }
  WM_REQUIRE(2 * L.g.cs < (1ll << 31) && 2 * Lc.g.cs < (1ll << 31),
             "expand3d: level too large for 32-bit indexing");
  // elements j = jlo..jlo+nj-1 hold the fine points ix = 2j-1, 2j in range
  int jlo = (ax.lo + 1) >> 1, nj = ((ax.hi + 1) >> 1) - jlo + 1;
  WM_REQUIRE(jlo + nj <= L.g.W, "expand3d: loop range outside the padded rows");
  int nyr = ay.hi - ay.lo + 1, nzr = az.hi - az.lo + 1;
  int bpc = (int)cdiv((long long)nyr * nj, 256);
  // 16 planes per thread, 4 when that leaves fewer than 1024 blocks
  const uint32_t *imask = fusepin ? L.imask.p : nullptr;
  if ((long long)bpc * cdiv(nzr, 16) >= 1024) {
    hipLaunchKernelGGL((k_expand_rb<S, SC, 16>), dim3(bpc * (int)cdiv(nzr, 16)), dim3(256), 0, s,
                       phi_of<S>(L).p, L.g, (const SC *)phi_of<SC>(Lc).p, Lc.g, ax, ay, az, jlo,
                       nj, bpc, imask, skip);
  } else {
    hipLaunchKernelGGL((k_expand_rb<S, SC, 4>), dim3(bpc * (int)cdiv(nzr, 4)), dim3(256), 0, s,
                       phi_of<S>(L).p, L.g, (const SC *)phi_of<SC>(Lc).p, Lc.g, ax, ay, az, jlo,
                       nj, bpc, imask, skip);
  }
  WM_LAUNCHED();
}

// residual3d + restrict3d as one kernel (k_resrestrict) where that is
// exact: six Dirichlet boundaries, plain 2:1 restriction windows on every
// axis, and no conductor record that zeroes a neighbour's residual.  The
// conductor points' residuals are written first by k_residual_cond.
static bool fused_restrict_ok(wm_mg3d *h, const Level &L, bool anyzero) {
  static const bool off = getenv("WARPMI_NO_FUSED_RESTRICT") != nullptr;
  if (off || anyzero) return false;
  for (int i = 0; i < 6; i++)
    if (h->bounds[i] != 0) return false;
  return L.rstd[0] && L.rstd[1] && L.rstd[2];
}

// Resident blocks per CU of k_resrestrict<S, TR, TC, MASK> by block size
// (index: wavefronts per block, 1..16; LDS not counted).  Filled by the
// first call, which wm_mg3d_create makes outside any stream capture.
template <typename S, typename TR, typename TC, bool MASK>
static const int *resrestrict_occupancy() {
  static int tab[17] = {-1};
  if (tab[0] < 0) {
    for (int k = 1; k <= 16; k++) {
      int n = 0;
      WM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &n, reinterpret_cast<const void *>(&k_resrestrict<S, TR, TC, MASK>), 64 * k, 0));
      tab[k] = std::max(n, 1);
    }
    tab[0] = 0;
  }
  return tab;
}

static void prepare_occupancy() {
  resrestrict_occupancy<double, double, double, false>();
  resrestrict_occupancy<double, double, double, true>();
  resrestrict_occupancy<float, float, float, false>();
  resrestrict_occupancy<float, float, float, true>();
  resrestrict_occupancy<bf16, float, float, false>();
  resrestrict_occupancy<bf16, float, float, true>();
}

// z chunk of a z-marching launch over `cols` columns and `ncz` coarse
// planes: each block walks 2 zc + 1 fine planes, and `slots` blocks run at
// once, so the time goes as ceil(blocks / slots) * (2 zc + 1).  Takes the
// cheapest zc up to 32, the larger one on a tie (fewer halo planes).
static int zchunk_for_rounds(long long cols, int ncz, long long slots) {
  int best = 1;
  long long bestcost = -1;
  for (int zc = 1; zc <= std::min(32, ncz); zc++) {
    long long nb = cols * cdiv(ncz, zc);
    long long cost = cdiv(nb, slots) * (2 * zc + 1);
    if (bestcost < 0 || cost <= bestcost) {
      bestcost = cost;
      best = zc;
    }
  }
  return best;
}

template <typename S, typename TR, typename TC>
static void resrestrict(wm_mg3d *h, Level &L, Level &Lc, int level, const Params &P,
                        const int *skip, hipStream_t s, TC *zfill, long long nzfill) {
  typedef typename St<S>::T T;
  typedef typename RestrictArith<TR, TC>::type RA;
  const S *phi = phi_of<S>(L).p;
  const S *rho = rho_of<S>(L).p;
  TR *res = res_of<TR>(L).p;
  double cst = 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
  T dxsqic = (T)(L.dxsqi * P.mgparam * cst), dysqic = (T)(L.dysqi * P.mgparam * cst),
    dzsqic = (T)(L.dzsqi * P.mgparam * cst);
  if (L.hasmask) {
    // the interior points' zeros are written once per solve
    // (zero_interior_res); here the records of both colours
    CondDev<T> &C = cond_of<T>(L);
    SubList<T> s0 = C.res[0][0].view(), s1 = C.res[0][1].view();
    if (s0.n + s1.n > 0) {
      hipLaunchKernelGGL((k_residual_rec2<S, TR, T, RBIdx>), dim3(blocks(s0.n + s1.n)), dim3(256),
                         0, s, phi, rho, res, RBIdx{L.g}, L.g, s0, s1, dxsqic, dysqic, dzsqic,
                         (T)P.mgparam, skip, rec_order(C.resord[0], s0.n + s1.n));
      WM_LAUNCHED();
    }
  }
  const int ncx = Lc.nx - 1, ncy = Lc.ny - 1, ncz = Lc.nz - 1;
  // tiles: up to 127 coarse columns and `maxpairs` fine pairs (one per
  // thread) per block, sizes balanced over the range (1024: 127 x 3 coarse
  // points, 896 threads at the finest level of 256x256x512, ~1% faster per
  // solve than 640); the z chunk from the kernel's occupancy
  // (zchunk_for_rounds)
  static const int mp_env = getenv("WARPMI_RESTRICT_PAIRS") ? atoi(getenv("WARPMI_RESTRICT_PAIRS")) : 0;
  const int maxpairs = mp_env >= 64 && mp_env <= 1024 ? mp_env : 1024;
  int ntx = (int)cdiv(ncx, 127);
  int cxt = (int)cdiv(ncx, ntx);
  int cytmax = std::max(1, (maxpairs / (cxt + 1) - 1) / 2);
  int nty = (int)cdiv(ncy, cytmax);
  int cyt = (int)cdiv(ncy, nty);
  long long cols = (long long)ntx * nty;
  int npairs = (cxt + 1) * (2 * cyt + 1);
  int threads = (int)(64 * cdiv(npairs, 64));
  size_t shm = 2 * sizeof(RA) * 2 * (cxt + 1) * (2 * cyt + 1);
  WM_REQUIRE(threads <= 1024 && shm <= 65536, "restrict3d: tile does not fit the fused kernel");
  const int *occ = L.hasmask ? resrestrict_occupancy<S, TR, TC, true>()
                             : resrestrict_occupancy<S, TR, TC, false>();
  long long per_cu = std::min<long long>(occ[threads / 64], (long long)(h->lds_per_cu / shm));
  int zc = zchunk_for_rounds(cols, ncz, std::max<long long>(1, per_cu) * h->ncu);
  static const int zc_env = getenv("WARPMI_RESTRICT_ZC") ? atoi(getenv("WARPMI_RESTRICT_ZC")) : 0;
  if (zc_env > 0) zc = zc_env;
  int ntz = (int)cdiv(ncz, zc);
  long long nb = cols * ntz;
  // extra blocks zeroing zfill, about 8 elements per thread
  long long nbz = nzfill > 0 ? std::max<long long>(1, cdiv(nzfill, 8ll * threads)) : 0;
  WM_REQUIRE(nb + nbz < (1ll << 31), "restrict3d: too many blocks");
  if (L.hasmask)
    hipLaunchKernelGGL((k_resrestrict<S, TR, TC, true>), dim3((int)(nb + nbz)), dim3(threads), shm,
                       s, phi, rho, (const TR *)res, rho_of<TC>(Lc).p, L.g, Lc.g, dxsqic, dysqic,
                       dzsqic, (T)P.mgparam, (const uint32_t *)L.mask.p, cxt, cyt, zc, ntx, nty,
                       (int)nb, L.ff, zfill, nzfill, skip);
  else
    hipLaunchKernelGGL((k_resrestrict<S, TR, TC, false>), dim3((int)(nb + nbz)), dim3(threads),
                       shm, s, phi, rho, (const TR *)res, rho_of<TC>(Lc).p, L.g, Lc.g, dxsqic,
                       dysqic, dzsqic, (T)P.mgparam, (const uint32_t *)nullptr, cxt, cyt, zc, ntx,
                       nty, (int)nb, L.ff, zfill, nzfill, skip);
  WM_LAUNCHED();
}

// Zeros of the interior conductor points in the residual arrays of the
// levels that use k_resrestrict (cond_potmgres, f3d_mgrid.F:2474-2506).  On
// those levels nothing else writes these points of res, so once per solve
// is enough; the records written after them in every V-cycle still win
// where one sits on an interior point, as in the reference.
static Params P2of(Params P) {
  P.mgform = 2;
  return P;
}

template <typename T>
static void zero_interior_res(wm_mg3d *h, const Params &P, hipStream_t s) {
  for (size_t l = 0; l + 1 < h->lv.size(); l++) {
    Level &L = *h->lv[l];
    if (!L.coarser || !L.hasmask || L.nint == 0) continue;
    CondDev<T> &C = cond_of<T>(L);
    if (!fused_restrict_ok(h, L, C.res[0][0].anyzero || C.res[0][1].anyzero)) continue;
    SubList<T> none{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL((k_residual_cond<T, T, T, RBIdx>), dim3(blocks(L.nint)), dim3(256), 0, s,
                       (const T *)nullptr, (const T *)nullptr, res_of<T>(L).p, RBIdx{L.g}, L.g,
                       L.nint, L.int_pos.p, none, (T)0., (T)0., (T)0., (T)P.mgparam, nullptr);
    WM_LAUNCHED();
  }
}

// vcycle (f3d_mgrid.F:1101-1591) on level data of storage type S (the
// coarser levels use Coarse<S>).
// zeroed: the caller has just filled this level's phi with zeros.  The pin
// at the start of the first pass then writes zeros over zeros (interior
// points are pinned to 0 unless on level 0 with mgform 1, which is never
// zero-filled) and the Neumann mirrors copy zeros, so it is left out.
template <typename S>
static void vcycle(wm_mg3d *h, int level, const Params &P, const int *skip, hipStream_t s,
                   bool zeroed = false, bool repin = true, double *cpart = nullptr,
                   double *cpart_end = nullptr, int *nslots = nullptr, bool pp = false);

// Exchanges the storage of the finest level's FP64 phi and save buffers.
static void pp_swap(wm_mg3d *h) {
  h->lv[0]->phi64.swap(h->save64);
  h->orient ^= 1;
}

// Whether the plain FP64 V-cycles that carry the fused convergence test
// keep their starting phi in the save buffer (exchanging the buffers)
// instead of copying every swept point there: the finest level runs half
// sweeps and at least one down pass (WARPMI_SAVE_SWAP=0 switches it off).
static bool pp_applies(const wm_mg3d *h, const Params &P) {
  static const bool off = getenv("WARPMI_SAVE_SWAP") && atoi(getenv("WARPMI_SAVE_SWAP")) == 0;
  static const bool tiled_env = getenv("WARPMI_SOR_TILED") != nullptr;
  return !off && !tiled_env && P.down >= 1 && P.up >= 1 && !fused_ok(h, *h->lv[0], 0, P);
}

// The FP64 V-cycle's tail (k_vtail) for the parameters P: the first level
// from 1 on with at most tail_nodes nodes, and every level from there on
// using the half sweeps (the fused pass swaps buffers, which the table does
// not follow).  Writes the table (synchronously: no V-cycle of this solver
// is in flight between solves) when its contents change.
static void prepare_tail(wm_mg3d *h, const Params &P) {
  h->tail.on = false;
  const int nlev = (int)h->lv.size();
  // the first level from which the levels' FP64 phi, rho and res fit the
  // budget (the tail starts there or below, once its tables fit too); the
  // FP32/BF16 V-cycles run these levels in FP64, tail or not, so that
  // every SOR form gives the same bits
  h->dc_double_from = -1;
  if (h->tail_lds == 0) return;
  {
    long long need = 0;
    for (int l = nlev - 1; l >= 1; l--) {
      const Level &L = *h->lv[l];
      need += (L.coarser ? 3 : 2) * 2 * L.g.cs;
      if ((size_t)need * sizeof(double) > h->tail_lds) break;
      h->dc_double_from = l;
    }
  }
  for (int l = 1; l < nlev; l++)
    if (fused_ok(h, *h->lv[l], l, P) || !h->lv[l]->phi64.p || !h->lv[l]->rho64.p ||
        (h->lv[l]->coarser && !h->lv[l]->res64.p))
      return;
  const int *b = h->bounds;
  const bool periodic = b[0] == 2 || b[2] == 2 || b[4] == 2;
  const SubList<double> none{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // The tables of the levels from `start` on and their gather list; the
  // start moves up a level at a time until everything fits the LDS budget
  // and the item list.
  for (int start = 1; start < nlev; start++) {
    std::vector<TailLevel> tab;
    std::vector<TailItem> items;
    int off = 0, gw = 0;
    // an LDS region of `words` (even offset); gathered from src when given
    auto region = [&](const void *src, long long words) -> int {
      const int o = off;
      off += (int)((words + 1) & ~1ll);
      if (src && words > 0) {
        items.push_back(TailItem{(const uint32_t *)src, o, (int)words, gw, 0});
        gw += (int)((words + TAIL_GB - 1) / TAIL_GB * TAIL_GB);
      }
      return o;
    };
    auto sub = [&](const SubList<double> &v, bool residual) {
      TailSub t{v.n, 0, 0, 0, 0, 0, 0};
      if (v.n == 0) return t;
      t.pos = region(v.pos, 3ll * v.n);
      t.modes = region(v.modes, v.n);
      t.k = region(v.k, 12ll * v.n);
      t.denom = region(v.denom, 2ll * v.n);
      if (residual) {
        t.fac = region(v.fac, 2ll * v.n);
        t.zero = region(v.zero, v.n);
      }
      return t;
    };
    for (int l = start; l < nlev; l++) {
      Level &L = *h->lv[l];
      TailLevel t;
      std::memset(&t, 0, sizeof(t));
      t.g = L.g;
      t.phi = L.phi64.p;
      const long long n2 = 2 * L.g.cs;
      // level arrays: only the first level's rho comes from HBM (its phi is
      // zero, the others' phi and rho are written by the restriction, and
      // the points of res that no residual writes are zero)
      t.lphi = region(nullptr, 2 * n2);
      t.lrho = region(l == start ? (const void *)L.rho64.p : nullptr, 2 * n2);
      if (L.coarser) t.lres = region(nullptr, 2 * n2);
      LoopRange sr = sor_range(L, b), rr = res_range(L, b);
      const int srv[6] = {sr.ixmin, sr.ixmax, sr.iymin, sr.iymax, sr.izmin, sr.izmax};
      const int rrv[6] = {rr.ixmin, rr.ixmax, rr.iymin, rr.iymax, rr.izmin, rr.izmax};
      for (int i = 0; i < 6; i++) {
        t.sr[i] = srv[i];
        t.rr[i] = rrv[i];
      }
      // padded element and row counts of the ranges (powers of two)
      auto lg2 = [](int n) {
        int k = 0;
        while ((1 << k) < n) k++;
        return k;
      };
      t.ssh[0] = lg2((sr.ixmax + 1) / 2 - (sr.ixmin + 1) / 2 + 1);
      t.ssh[1] = lg2(sr.iymax - sr.iymin + 1);
      t.rsh[0] = lg2((rr.ixmax + 1) / 2 - (rr.ixmin + 1) / 2 + 1);
      t.rsh[1] = lg2(rr.iymax - rr.iymin + 1);
      // the coefficients as halfpass and residual_rb form them
      const double cst = P.mgparam * 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
      t.sc[0] = L.dxsqi * cst;
      t.sc[1] = L.dysqi * cst;
      t.sc[2] = L.dzsqi * cst;
      t.sc[3] = 1. - P.mgparam;
      const double cr = 0.5 / (L.dxsqi + L.dysqi + L.dzsqi);
      t.rc[0] = L.dxsqi * P.mgparam * cr;
      t.rc[1] = L.dysqi * P.mgparam * cr;
      t.rc[2] = L.dzsqi * P.mgparam * cr;
      t.rs[0] = P.mgparam * L.dxsqi * cr;
      t.rs[1] = P.mgparam * L.dysqi * cr;
      t.rs[2] = P.mgparam * L.dzsqi * cr;
      CondDev<double> &C = cond_of<double>(L);
      int f = 0;
      const bool mirror = P.icnd == 2 && L.neumann;
      if (L.sub_on_interior || periodic || mirror) f |= TAIL_PIN_AFTER;
      if (mirror) f |= TAIL_MIRROR;
      if (L.sub_on_interior && L.nint > 0) f |= TAIL_SPLIT;
      if (L.coarser && P.up > 0 && pin_fusable(h, L, l, P)) f |= TAIL_FUSEPIN;
      if (L.hasmask) f |= TAIL_HASMASK;
      if (!(L.sub_on_interior && L.nint > 0) && !C.res[0][0].anyzero && !C.res[0][1].anyzero)
        f |= TAIL_RES1;
      t.flags = f;
      if (L.hasmask) t.lmask = region(L.mask.p, (long long)L.mask.n);
      if (f & TAIL_FUSEPIN) t.limask = region(L.imask.p, (long long)L.imask.n);
      t.nint = L.nint;
      if (L.nint > 0) t.lint = region(L.int_pos.p, 3ll * L.nint);
      for (int c = 0; c < 2; c++) {
        t.sor[c] = sub(P.lcnd ? C.sor[c].view() : none, false);
        t.rsub[c] = sub(L.hasmask && L.coarser ? C.res[0][c].view() : none, true);
      }
      for (int p = 0; p < 2; p++) {
        t.nsub[p] = (int)(L.subpos_h[p].size() / 3);
        t.subpos[p] = L.subpos_d[p].p;
        t.subdels[p] = L.subdels_d[p].p;
      }
      if (L.coarser) {
        for (int a = 0; a < 3; a++) {
          t.ra[a] = TailAxisR{region(L.rlo[a].p, (long long)L.rlo[a].n),
                              region(L.rhi[a].p, (long long)L.rhi[a].n),
                              region(L.rw[a].p, 2ll * (long long)L.rw[a].n), L.rcmin[a],
                              L.rcmax[a]};
          t.ea[a] = TailAxisE{region(L.ej[a].p, (long long)L.ej[a].n),
                              region(L.ew[a].p, 2ll * (long long)L.ew[a].n), L.elo[a], L.ehi[a]};
        }
        t.ff = L.ff;
      }
      tab.push_back(t);
    }
    if ((size_t)off * 4 > h->tail_lds || (int)items.size() > TAIL_MAX_ITEMS) continue;
    // the steps, in the order vcycle and its callees launch them for these
    // levels (phi of the first level starts at zero)
    std::vector<TailOp> ops;
    const int nl = (int)tab.size();
    const bool anybc = b[0] || b[1] || b[2] || b[3] || b[4] || b[5];
    auto pass = [&](int l, bool pin) {
      const TailLevel &t = tab[l];
      const bool sweeps = t.sr[3] >= t.sr[2] && t.sr[5] >= t.sr[4];
      if (pin) ops.push_back(TailOp{TOP_PIN, l, 0, 0});
      for (int c = 0; c < 2; c++) {
        ops.push_back(TailOp{TOP_HALF, l, c, 0});
        if (sweeps && (t.flags & TAIL_PIN_AFTER)) ops.push_back(TailOp{TOP_PIN, l, 0, 0});
        if (anybc) ops.push_back(TailOp{TOP_BCPHI, l, 0, 0});
      }
    };
    for (int l = 0; l < nl; l++) {
      for (int i = 0; i < P.down; i++) pass(l, i > 0);
      if (l + 1 < nl) {
        ops.push_back(TailOp{TOP_RESID, l, 0, 0});
        ops.push_back(TailOp{TOP_RESTRICT, l, 0, 0});
      }
    }
    for (int l = nl - 1; l >= 0; l--) {
      bool fusepin = false;
      if (l + 1 < nl) {
        ops.push_back(TailOp{TOP_EXPAND, l, 0, 0});
        fusepin = tab[l].flags & TAIL_FUSEPIN;
      }
      for (int i = 0; i < P.up; i++) pass(l, !(fusepin && i == 0));
    }
    if ((int)ops.size() > TAIL_MAX_OPS) return;
    TailRun run{};
    run.nops = (int)ops.size();
    for (int i = 0; i < 6; i++) run.b[i] = b[i];
    run.mgparam = P.mgparam;
    run.lds_words = off;
    run.nitems = (int)items.size();
    run.gather_words = gw;
    if (h->tailtab.n < (size_t)nlev) h->tailtab.alloc((size_t)nlev);
    if (h->tailitems.n < (size_t)TAIL_MAX_ITEMS) h->tailitems.alloc((size_t)TAIL_MAX_ITEMS);
    if (h->tailops.n < (size_t)TAIL_MAX_OPS) h->tailops.alloc((size_t)TAIL_MAX_OPS);
    const bool same =
        h->tail.host.size() == tab.size() && h->tail.items.size() == items.size() &&
        h->tail.ops.size() == ops.size() &&
        std::memcmp(h->tail.host.data(), tab.data(), sizeof(TailLevel) * tab.size()) == 0 &&
        (items.empty() ||
         std::memcmp(h->tail.items.data(), items.data(), sizeof(TailItem) * items.size()) == 0) &&
        (ops.empty() || std::memcmp(h->tail.ops.data(), ops.data(), sizeof(TailOp) * ops.size()) == 0);
    if (!same) {
      // captured V-cycles hold the old step count and gather sizes by value
      h->drop_graphs();
      h->gkey = wm_mg3d::GKey();
      WM_HIP(hipStreamSynchronize(h->s));
      WM_HIP(hipMemcpy(h->tailtab.p, tab.data(), sizeof(TailLevel) * tab.size(),
                       hipMemcpyHostToDevice));
      if (!items.empty())
        WM_HIP(hipMemcpy(h->tailitems.p, items.data(), sizeof(TailItem) * items.size(),
                         hipMemcpyHostToDevice));
      if (!ops.empty())
        WM_HIP(hipMemcpy(h->tailops.p, ops.data(), sizeof(TailOp) * ops.size(),
                         hipMemcpyHostToDevice));
      h->tail.host = tab;
      h->tail.items = items;
      h->tail.ops = ops;
    }
    h->tail.down = P.down;
    h->tail.up = P.up;
    h->tail.run = run;
    h->tail.start = start;
    h->tail.on = true;
    return;
  }
}

// Whether vcycle<double> on this level runs as the tail launch for P.
static bool tail_applies(const wm_mg3d *h, int level, const Params &P) {
  return h->tail.on && level == h->tail.start && h->tail.down == P.down &&
         h->tail.up == P.up && h->tail.run.mgparam == P.mgparam;
}

// The coarse-grid correction of one level (f3d_mgrid.F:1293-1507): the
// residual restricted to the next level (stored as SC), the V-cycle there
// from zero, the expansion added back.  Returns whether the expansion also
// did the pin of the first up pass.
template <typename S, typename SC>
static bool coarse_correction(wm_mg3d *h, Level &L, int level, const Params &P,
                              const int *skip, hipStream_t s) {
  typedef typename St<S>::T T;
  Level &Lc = *h->lv[level + 1];
  bool anyzero = false;
  if (L.hasmask) {
    ensure_records<T>(h, L, level, P, true, 0);
    anyzero = cond_of<T>(L).res[0][0].anyzero || cond_of<T>(L).res[0][1].anyzero;
  }
  // the coarse phi starts at zero: zeroed by the fused kernel's extra
  // blocks, or by k_fill after the separate residual and restriction
  if (fused_restrict_ok(h, L, anyzero)) {
    resrestrict<S, T, SC>(h, L, Lc, level, P, skip, s, phi_of<SC>(Lc).p, 2 * Lc.g.cs);
  } else {
    residual_rb<S, T>(h, L, level, P, phi_of<S>(L).p, rho_of<S>(L).p, res_of<T>(L).p, 0, skip,
                      s);
    restrict_to<T, SC>(L, (const T *)res_of<T>(L).p, RBIdx{L.g}, rho_of<SC>(Lc).p,
                       RBIdx{Lc.g}, skip, s, Lc.nx, Lc.ny, Lc.nz);
    hipLaunchKernelGGL((k_fill<SC>), dim3(blocks(2 * Lc.g.cs)), dim3(256), 0, s,
                       phi_of<SC>(Lc).p, 2 * Lc.g.cs, (SC)0., skip);
    WM_LAUNCHED();
  }
  vcycle<SC>(h, level + 1, P, skip, s, true);
  const bool fusepin = P.up > 0 && pin_fusable(h, L, level, P);
  expand_rb<S, SC>(h, L, Lc, level, P, skip, s, fusepin);
  launch_bc<S>(phi_of<S>(L).p, RBIdx{L.g}, L.nx, L.ny, L.nz, h->bounds, 0, 0, skip, s);
  return fusepin;
}

template <typename S>
static void vcycle(wm_mg3d *h, int level, const Params &P, const int *skip, hipStream_t s,
                   bool zeroed, bool repin, double *cpart, double *cpart_end, int *nslots,
                   bool pp) {
  typedef typename Coarse<S>::type SC;
  Level &L = *h->lv[level];
  // pp (finest level, FP64): the V-cycle's starting phi moves to the save
  // buffer by exchanging the two buffers, the first pass sweeps from it
  // into phi, and the convergence test compares with it without copying
  const S *src = nullptr;
  if constexpr (std::is_same<S, double>::value) {
    if (pp) {
      WM_REQUIRE(level == 0 && !repin && P.down >= 1 && cpart,
                 "vcycle: the two-buffer form needs the finest level, no leading pin, a down "
                 "pass and the fused convergence test");
      pp_swap(h);
      src = h->save64.p;
    }
  }
  const bool novolt = !(level == 0 && P.mgform == 1);
  bool fusepin = false;
  if constexpr (std::is_same<S, double>::value) {
    // levels from the tail's start on, entered with phi at zero
    if (zeroed && level > 0 && tail_applies(h, level, P)) {
      hipLaunchKernelGGL(k_vtail, dim3(1), dim3(1024), sizeof(uint32_t) * h->tail.run.lds_words,
                         s, (const TailLevel *)h->tailtab.p, (const TailItem *)h->tailitems.p,
                         (const TailOp *)h->tailops.p, h->tail.run, skip);
      WM_LAUNCHED();
      return;
    }
  }
  // The first pass pins unless phi starts at zero without voltages, or
  // (repin false) the caller knows the pinned points still hold their
  // values from the previous V-cycle.
  for (int i = 0; i < P.down; i++)
    sorpass<S>(h, L, level, P, skip, s, !(zeroed && novolt && i == 0) && (repin || i > 0),
               nullptr, nullptr, nullptr, i == 0 ? src : nullptr);
  if (L.coarser) {
    // the coarsest level of a lower-precision V-cycle is solved in FP64,
    // and so are the levels the single-launch tail can hold
    if (!std::is_same<S, double>::value &&
        (level + 2 == (int)h->lv.size() || level + 1 == h->dc_double_from))
      fusepin = coarse_correction<S, double>(h, L, level, P, skip, s);
    else
      fusepin = coarse_correction<S, SC>(h, L, level, P, skip, s);
  }
  // the first up pass's pin already ran inside the expansion when fusepin;
  // the last up pass carries the convergence test when cpart is given
  for (int i = 0; i < P.up; i++)
    sorpass<S>(h, L, level, P, skip, s, !(fusepin && i == 0), i == P.up - 1 ? cpart : nullptr,
               cpart_end, nslots, nullptr, !pp);
}

template <typename T>
static void prepare_levels(wm_mg3d *h, const Params &P) {
  for (size_t l = 0; l < h->lv.size(); l++) {
    Level &L = *h->lv[l];
    ensure_buffers<T>(L, L.coarser);
    build_mask(h, L, (int)l, P);
    ensure_records<T>(h, L, (int)l, P, false, 0);
    if (L.coarser && L.hasmask) ensure_records<T>(h, L, (int)l, P, true, 0);
  }
}

static int nconv_blocks(const RBGeom &g) {
  long long rows = 2ll * g.PY * g.PZ;
  return (int)std::min<long long>(std::min(1024, WM_CONV_BLOCKS), std::max<long long>(1, cdiv(rows, 8)));
}

// The low-precision half of a defect-correction iteration: the FP64
// residual becomes the source of the error equation in storage type S, the
// V-cycle runs from a zero error, and the error is added to the FP64 phi.
template <typename S>
static void dc_lowp(wm_mg3d *h, const Params &Pinit, const double *base, long long np,
                    bool fused, double mgtol, int mgmaxiters, const int *skip, hipStream_t s) {
  Level &L0 = *h->lv[0];
  residual_rb<double, S>(h, L0, 0, Pinit, base, h->rhosave64.p, rho_of<S>(L0).p, 1, skip, s,
                         true);
  WM_REQUIRE(phi_of<S>(L0).n >= (size_t)np, "defect correction: error array too small");
  zero_buf<S>(phi_of<S>(L0), skip, s);
  vcycle<S>(h, 0, Pinit, skip, s, true);
  if (fused) {
    WM_KW_DISPATCH((L0.nx + 4) / 2, hipLaunchKernelGGL((k_dc_update<S, KW>), dim3(nconv_blocks(L0.g)),
                                              dim3(256), 0, s, L0.phi64.p,
                                              (const S *)phi_of<S>(L0).p, L0.g, h->state.p,
                                              h->partials.p));
    WM_LAUNCHED();
    hipLaunchKernelGGL(k_converge_close, dim3(1), dim3(1024), 0, s, h->partials.p,
                       nconv_blocks(L0.g), h->state.p, mgtol, mgmaxiters);
  } else {
    hipLaunchKernelGGL((k_dc_add<S>), dim3(blocks(np)), dim3(256), 0, s, L0.phi64.p,
                       (const S *)phi_of<S>(L0).p, h->save64.p, np, skip);
  }
  WM_LAUNCHED();
}

// One iteration of the multigrid3dsolve loop (f3d_mgrid.F:313-434).
// Whether a V-cycle of the plain FP64 solve may skip the pin that starts
// its first pass on the finest level: the previous V-cycle ended with
// passes that began with a pin (after the expansion; so there must be an
// up pass -- without one the expansion's changes to pinned points stand
// until the next V-cycle's leading pin), and under these conditions nothing
// since wrote a pinned point -- the dense sweeps skip interior points, no
// subgrid record writes one, no periodic wrap copies over one, no Neumann
// ghost needs mirroring (the same conditions under which halfpass leaves
// out the pin after a sweep); the convergence test only reads phi.
static bool repin_skippable(const wm_mg3d *h, const Params &P) {
  const Level &L0 = *h->lv[0];
  const int *b = h->bounds;
  bool periodic = b[0] == 2 || b[2] == 2 || b[4] == 2;
  return P.up >= 1 && !(L0.sub_on_interior || periodic || (P.icnd == 2 && L0.neumann));
}

static void iteration(wm_mg3d *h, int prec, const Params &P, double mgtol, int mgmaxiters,
                      hipStream_t s, bool repin = true) {
  Level &L0 = *h->lv[0];
  const int *skip = &h->state.p->done;
  if (prec == WM_PREC_FP64 && P.mgform == 1) {
    // Where the leading pin is left out (repin false) nothing but the last
    // up pass's sweeps changes a point between its save at the end of the
    // previous V-cycle and the test: those sweeps compare every point they
    // write with save and copy it there, and k_converge is not needed.  The
    // first V-cycle of a solve (pin, full copy) keeps k_converge.  Where
    // pp_applies, the copy is left out instead: the V-cycle exchanges the
    // buffers, so that save holds its starting phi, and its first pass
    // sweeps from save into phi (every point a sweep writes, both colours;
    // the points no sweep writes -- boundaries, pinned interior points --
    // hold the same values in both buffers).
    static const bool tiled_env = getenv("WARPMI_SOR_TILED") != nullptr;
    const bool fuseconv = !repin && P.up >= 1 && h->convparts.p && !tiled_env;
    int nslots = 0;
    vcycle<double>(h, 0, P, skip, s, false, repin, fuseconv ? h->convparts.p : nullptr,
                   h->convparts.p + h->convparts.n, &nslots, fuseconv && pp_applies(h, P));
    if (fuseconv) {
      hipLaunchKernelGGL(k_converge_close, dim3(1), dim3(1024), 0, s, h->convparts.p, nslots,
                         h->state.p, mgtol, mgmaxiters);
      WM_LAUNCHED();
      return;
    }
  } else {
    // Residual-correction form: the error equation on the residual of the
    // saved phi (mgform == 2), with the V-cycle in double or float.
    // Without periodic boundaries the update is done in place on phi64
    // and fused with the convergence test (k_dc_update); with them the
    // boundary conditions must come between the update and the test.
    const int *b = h->bounds;
    bool fused = prec != WM_PREC_FP64 && !(b[0] == 2 || b[2] == 2 || b[4] == 2);
    double *base = fused ? L0.phi64.p : h->save64.p;
    Params Pinit = P;
    Pinit.mgform = 2;
    // the pin changes phi only in the first iteration (the error is zero at
    // the interior points, so the update leaves their voltages); later
    // iterations skip it (its skip flag is the iteration count), unless
    // Neumann mirrors come with it
    const int *pinskip = (P.icnd == 2 && L0.neumann) ? skip : &h->state.p->iters;
    pin_and_mirror<double>(h, L0, 0, Pinit, base, 1, pinskip, s);
    long long np = 2 * L0.g.cs;
    if (prec == WM_PREC_FP64) {
      residual_rb<double, double>(h, L0, 0, Pinit, base, h->rhosave64.p, L0.rho64.p, 1, skip, s,
                                  true);
      zero_buf<double>(L0.phi64, skip, s);
      // the error starts at zero in both buffers (their boundary values
      // must agree for the fused pass)
      zero_buf<double>(L0.phi64b, skip, s);
      vcycle<double>(h, 0, Pinit, skip, s, true);
      // phi = phi + phisave (the error plus the saved phi)
      hipLaunchKernelGGL((k_dc_add<double>), dim3(blocks(np)), dim3(256), 0, s, L0.phi64.p,
                         L0.phi64.p, h->save64.p, np, skip);
      WM_LAUNCHED();
    } else if (prec == WM_PREC_FP32_DC) {
      dc_lowp<float>(h, Pinit, base, np, fused, mgtol, mgmaxiters, skip, s);
    } else {
      dc_lowp<bf16>(h, Pinit, base, np, fused, mgtol, mgmaxiters, skip, s);
    }
    launch_bc<double>(L0.phi64.p, RBIdx{L0.g}, L0.nx, L0.ny, L0.nz, h->bounds, 0, 0, skip, s);
    if (fused) return;
  }
  WM_KW_DISPATCH((L0.nx + 4) / 2, hipLaunchKernelGGL((k_converge<double, KW>), dim3(nconv_blocks(L0.g)),
                                            dim3(256), 0, s, L0.phi64.p, h->save64.p, L0.g,
                                            h->state.p, h->partials.p, 1));
  WM_LAUNCHED();
  hipLaunchKernelGGL(k_converge_close, dim3(1), dim3(1024), 0, s, h->partials.p,
                     nconv_blocks(L0.g), h->state.p, mgtol, mgmaxiters);
  WM_LAUNCHED();
}

__global__ void k_state_reset(SolveState *st) {
  st->iters = 0;
  st->done = 0;
  st->err = 0.;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int wm_mg3d_create(wm_mg3d **out, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                   double dz, const int64_t *bounds, int64_t mgmaxlevels) {
  WM_ENTRY({
    WM_REQUIRE(out != nullptr, "wm_mg3d_create: out is NULL");
    WM_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "wm_mg3d_create: need nx, ny, nz >= 2");
    WM_REQUIRE(dx > 0 && dy > 0 && dz > 0, "wm_mg3d_create: cell sizes must be positive");
    WM_REQUIRE(nx < (1 << 30) && ny < (1 << 30) && nz < (1 << 30) &&
                   2 * make_geom((int)nx, (int)ny, (int)nz).cs + 64 < (1ll << 31),
               "wm_mg3d_create: grid too large (the red-black arrays are indexed with 32 bits)");
    for (int i = 0; i < 6; i++)
      WM_REQUIRE(bounds[i] >= 0 && bounds[i] <= 2,
                 "wm_mg3d_create: bounds must be 0 (Dirichlet), 1 (Neumann) or 2 (periodic)");
    for (int a = 0; a < 3; a++)
      WM_REQUIRE((bounds[2 * a] == 2) == (bounds[2 * a + 1] == 2),
                 "wm_mg3d_create: periodic bounds must be set on both sides of an axis");
    std::unique_ptr<wm_mg3d> h(new wm_mg3d());
    h->nx = nx;
    h->ny = ny;
    h->nz = nz;
    h->dx = dx;
    h->dy = dy;
    h->dz = dz;
    for (int i = 0; i < 6; i++) h->bounds[i] = (int)bounds[i];
    // Levels: getmglevels (f3d_mgrid.F:3654-3845), capped as vcycle does.
    int64_t lnx = nx, lny = ny, lnz = nz;
    double ldx = dx, ldy = dy, ldz = dz;
    for (int level = 0;; level++) {
      std::unique_ptr<Level> L(new Level());
      L->nx = (int)lnx;
      L->ny = (int)lny;
      L->nz = (int)lnz;
      L->dx = ldx;
      L->dy = ldy;
      L->dz = ldz;
      L->dxsqi = 1. / (ldx * ldx);
      L->dysqi = 1. / (ldy * ldy);
      L->dzsqi = 1. / (ldz * ldz);
      L->g = make_geom((int)lnx, (int)lny, (int)lnz);
      bool more = (lnx >= 4 && lny >= 4 && lnz >= 4 && level < mgmaxlevels && level < 100);
      h->lv.push_back(std::move(L));
      if (!more) break;
      int64_t nxc, nyc, nzc;
      double dxc, dyc, dzc;
      next_coarse(lnx, lny, lnz, ldx, ldy, ldz, nxc, nyc, nzc, dxc, dyc, dzc);
      Level &F = *h->lv.back();
      F.coarser = true;
      F.ff = (F.dxsqi + F.dysqi + F.dzsqi) /
             (1. / (dxc * dxc) + 1. / (dyc * dyc) + 1. / (dzc * dzc));
      restrict_axis((int)lnx, (int)nxc, h->bounds[0], h->bounds[1], F, 0);
      restrict_axis((int)lny, (int)nyc, h->bounds[2], h->bounds[3], F, 1);
      restrict_axis((int)lnz, (int)nzc, h->bounds[4], h->bounds[5], F, 2);
      expand_axis((int)lnx, (int)nxc, h->bounds[0], h->bounds[1], F, 0);
      expand_axis((int)lny, (int)nyc, h->bounds[2], h->bounds[3], F, 1);
      expand_axis((int)lnz, (int)nzc, h->bounds[4], h->bounds[5], F, 2);
      lnx = nxc;
      lny = nyc;
      lnz = nzc;
      ldx = dxc;
      ldy = dyc;
      ldz = dzc;
    }
    h->state.alloc(1);
    h->partials.alloc(WM_CONV_BLOCKS);
    WM_HIP(hipHostMalloc((void **)&h->state_h, sizeof(SolveState), hipHostMallocDefault));
    WM_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    int dev = 0, ncu = 0, lds = 0;
    WM_HIP(hipGetDevice(&dev));
    WM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    WM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    if (ncu > 0) h->ncu = ncu;
    if (lds > 0) h->lds_per_cu = (size_t)lds;
    // LDS of the V-cycle tail: what one workgroup may hold, less the item
    // list and 1 KB
    {
      int blk = 0;
      WM_HIP(hipDeviceGetAttribute(&blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
      const size_t fixed = sizeof(TailItem) * TAIL_MAX_ITEMS + 1024;
      size_t cap = (size_t)blk > fixed ? (size_t)blk - fixed : 0;
      if (getenv("WARPMI_TAIL_LDS")) cap = std::min<size_t>(cap, (size_t)atoll(getenv("WARPMI_TAIL_LDS")));
      if (cap > 65536 && hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vtail),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)cap) != hipSuccess) {
        (void)hipGetLastError();
        cap = 65536;
      }
      h->tail_lds = cap;
    }
    prepare_occupancy();
    WM_HIP(hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming));
    WM_HIP(hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming));
    *out = h.release();
  });
}

int wm_mg3d_destroy(wm_mg3d *h) {
  WM_ENTRY({
    if (h) {
      (void)hipDeviceSynchronize();
      delete h;
    }
  });
}

int wm_mg3d_nlevels(const wm_mg3d *h, int64_t *n) {
  WM_ENTRY({
    WM_REQUIRE(h && n, "wm_mg3d_nlevels: NULL argument");
    *n = (int64_t)h->lv.size();
  });
}

int wm_mg3d_level_dims(const wm_mg3d *h, int64_t level, int64_t *nxyz, double *dxyz) {
  WM_ENTRY({
    WM_REQUIRE(h && level >= 0 && level < (int64_t)h->lv.size(), "wm_mg3d_level_dims: bad level");
    const Level &L = *h->lv[level];
    if (nxyz) {
      nxyz[0] = L.nx;
      nxyz[1] = L.ny;
      nxyz[2] = L.nz;
    }
    if (dxyz) {
      dxyz[0] = L.dx;
      dxyz[1] = L.dy;
      dxyz[2] = L.dz;
    }
  });
}

int wm_mg3d_set_conductors(wm_mg3d *h, const wm_conductor_lists *c) {
  WM_ENTRY({
    WM_REQUIRE(h != nullptr, "wm_mg3d_set_conductors: NULL handle");
    for (size_t l = 0; l < h->lv.size(); l++) {
      Level &L = *h->lv[l];
      L.nint = 0;
      L.int_pos.release();
      L.int_volt.release();
      for (int p = 0; p < 2; p++) {
        L.subpos_h[p].clear();
        L.subdels_h[p].clear();
        L.subvolt_h[p].clear();
        L.subpos_d[p].release();
        L.subdels_d[p].release();
      }
      L.neumann = false;
      L.hasmask = false;
      L.mask.release();
      L.imask.release();
      L.c64.reset();
      L.c32.reset();
    }
    h->conductors_set = false;
    h->last_ok = false;
    h->drop_graphs();
    h->gkey = wm_mg3d::GKey();
    if (c == nullptr) return 0;
    WM_REQUIRE(c->levels >= (int64_t)h->lv.size() || (c->int_n == 0 && c->sub_n[0] == 0 && c->sub_n[1] == 0),
               "wm_mg3d_set_conductors: conductor data has fewer levels than the solver");
    for (size_t l = 0; l < h->lv.size(); l++) {
      Level &L = *h->lv[l];
      if (c->int_n > 0) {
        int64_t a = c->int_istart[l], b = c->int_istart[l + 1];
        std::vector<int> pos;
        std::vector<double> volt;
        for (int64_t i = a; i < b; i++) {
          int ix = (int)c->int_indx[3 * i], iy = (int)c->int_indx[3 * i + 1],
              iz = (int)c->int_indx[3 * i + 2];
          WM_REQUIRE(ix >= 0 && ix <= L.nx && iy >= 0 && iy <= L.ny && iz >= 0 && iz <= L.nz,
                     "wm_mg3d_set_conductors: interior point outside the level's grid");
          pos.insert(pos.end(), {ix, iy, iz});
          volt.push_back(c->int_volt[i]);
        }
        L.nint = (int)volt.size();
        L.int_pos.upload(pos);
        L.int_volt.upload(volt);
      }
      for (int p = 0; p < 2; p++) {
        if (c->sub_n[p] == 0) continue;
        int64_t a = c->sub_istart[p][l], b = c->sub_istart[p][l + 1];
        for (int64_t i = a; i < b; i++) {
          int ix = (int)c->sub_indx[p][3 * i], iy = (int)c->sub_indx[p][3 * i + 1],
              iz = (int)c->sub_indx[p][3 * i + 2];
          WM_REQUIRE(ix >= 0 && ix <= L.nx && iy >= 0 && iy <= L.ny && iz >= 0 && iz <= L.nz,
                     "wm_mg3d_set_conductors: subgrid point outside the level's grid");
          WM_REQUIRE(((ix + iy + iz) & 1) == p,
                     "wm_mg3d_set_conductors: subgrid point in the wrong parity list");
          L.subpos_h[p].insert(L.subpos_h[p].end(), {ix, iy, iz});
          for (int k = 0; k < 6; k++) {
            double d = c->sub_dels[p][6 * i + k];
            L.subdels_h[p].push_back(d);
            L.subvolt_h[p].push_back(c->sub_volt[p][6 * i + k]);
            if (0. >= d && d > -1.) L.neumann = true;
          }
        }
        L.subpos_d[p].upload(L.subpos_h[p]);
        L.subdels_d[p].upload(L.subdels_h[p]);
      }
    }
    h->conductors_set = true;
  });
}

int wm_mg3d_solve(wm_mg3d *h, double *phi, double *rho, double mgparam, int64_t mgform,
                  int64_t mgmaxiters, double mgtol, int64_t downpasses, int64_t uppasses,
                  int64_t lcndbndy, int64_t icndbndy, int64_t precision, int64_t *mgiters,
                  double *mgerror, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(h && phi && rho && mgiters && mgerror, "wm_mg3d_solve: NULL argument");
    WM_REQUIRE(mgform == 1 || mgform == 2, "wm_mg3d_solve: mgform must be 1 or 2");
    WM_REQUIRE(icndbndy == 1 || icndbndy == 2, "wm_mg3d_solve: icndbndy must be 1 or 2");
    WM_REQUIRE(precision == WM_PREC_FP64 || precision == WM_PREC_FP32_DC ||
                   precision == WM_PREC_BF16_DC,
               "wm_mg3d_solve: unknown precision");
    WM_REQUIRE(downpasses >= 0 && uppasses >= 0, "wm_mg3d_solve: negative pass count");
    int prec = (int)precision;
    hipStream_t us = wm_stream(stream);
    hipStream_t s = h->s;
    Params P{mgparam, (int)mgform, (int)downpasses, (int)uppasses, (int)lcndbndy, (int)icndbndy};
    Level &L0 = *h->lv[0];
    double eps0 = wm_eps0();
    double reps0c = mgparam / (eps0 * 2. * (L0.dxsqi + L0.dysqi + L0.dzsqi));
    double rinv = 1. / reps0c;
    // Buffers and conductor records.
    prepare_levels<double>(h, P);
    bool dc = !(prec == WM_PREC_FP64 && P.mgform == 1);
    if (prec != WM_PREC_FP64)
      prepare_levels<float>(h, Params{P.mgparam, 2, P.down, P.up, P.lcnd, P.icnd});
    if (prec == WM_PREC_BF16_DC) {
      L0.phi16.alloc_rb(2 * L0.g.cs);
      L0.phi16b.alloc_rb(2 * L0.g.cs);
      L0.rho16.alloc_rb(2 * L0.g.cs);
    }
    if (dc) {
      L0.res64.alloc_rb(std::max<size_t>(2 * L0.g.cs, (size_t)(L0.nx + 3) * (L0.ny + 3) * (L0.nz + 3)));
      h->rhosave64.alloc_rb(2 * L0.g.cs);
      Params P2 = P;
      P2.mgform = 2;
      if (prec == WM_PREC_FP64) prepare_levels<double>(h, P2);
      ensure_records<double>(h, L0, 0, P2, true, 1);
    }
    h->save64.alloc_rb(2 * L0.g.cs);
    prepare_tail(h, P);
    h->tail_used = h->tail.on && prec == WM_PREC_FP64 && mgmaxiters > 0;
    // Inputs.
    WM_HIP(hipEventRecord(h->ev_in, us));
    WM_HIP(hipStreamWaitEvent(s, h->ev_in, 0));
    long long nr = (long long)(h->nx + 1) * (h->ny + 1) * (h->nz + 1);
    long long nn = (long long)(h->nx + 3) * (h->ny + 3) * (h->nz + 3);
    long long np = 2 * L0.g.cs;
    double *rho0 = dc ? h->rhosave64.p : L0.rho64.p;
    hipLaunchKernelGGL(k_rho_in, dim3(blocks(nr)), dim3(256), 0, s, rho, rho0, L0.g, reps0c, rinv);
    WM_LAUNCHED();
    // The caller's phi is the one the last solve returned (reuse): phi64
    // and save64 still hold it.  With six Dirichlet bounds no kernel reads
    // phi64's guard cells and the output's guards are recomputed, so the
    // result is the same bits as after a fresh conversion.
    bool alldir = true;
    for (int i = 0; i < 6; i++) alldir = alldir && h->bounds[i] == 0;
    const bool fp64plain = prec == WM_PREC_FP64 && P.mgform == 1;
    // the V-cycles after the first keep their starting phi in save64
    // (exchanging the buffers); then save64 holds a stale phi at the end, so
    // a solve without that form cannot reuse the state of one with it
    const bool pp_now = fp64plain && repin_skippable(h, P) && pp_applies(h, P);
    const bool reuse = h->reuse_next && h->last_ok && phi == h->last_phi && fp64plain &&
                       alldir && mgmaxiters > 0 && repin_skippable(h, P) &&
                       !(h->last_pp && !pp_now);
    h->reuse_next = false;
    h->last_ok = false;
    h->last_reused = reuse;
    if (!reuse) {
      // the saved phi: written by the conversion itself when the boundary
      // conditions change nothing (six Dirichlet bounds), else copied after
      // the fused pass's second buffer needs the boundary values only when
      // the plain FP64 V-cycle may run that pass on the finest level
      // (the FP32/BF16 defect correction with six Dirichlet bounds updates
      // phi in place and tests the change itself: no saved phi)
      hipLaunchKernelGGL((k_nat_to_rb<double>), dim3(blocks(nn)), dim3(256), 0, s, phi,
                         L0.phi64.p, L0.g, alldir && prec == WM_PREC_FP64 ? h->save64.p : nullptr,
                         fp64plain && h->sor_form != 0 ? L0.phi64b.p : nullptr);
      WM_LAUNCHED();
      launch_bc<double>(L0.phi64.p, RBIdx{L0.g}, L0.nx, L0.ny, L0.nz, h->bounds, 0, 0, nullptr,
                        s);
      if (!alldir) {
        hipLaunchKernelGGL((k_copy<double, double>), dim3(blocks(np)), dim3(256), 0, s,
                           L0.phi64.p, h->save64.p, np, nullptr);
        WM_LAUNCHED();
      }
    }
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(1), 0, s, h->state.p);
    WM_LAUNCHED();
    if (prec == WM_PREC_FP64)
      zero_interior_res<double>(h, dc ? P2of(P) : P, s);
    else
      zero_interior_res<float>(h, Params{P.mgparam, 2, P.down, P.up, P.lcnd, P.icnd}, s);
    // the fixed zeros of the defect-correction source (its residual runs
    // in the steady form)
    if (dc) {
      if (prec == WM_PREC_FP64) residual_zeros<double>(h, L0, L0.rho64.p, s);
      else if (prec == WM_PREC_FP32_DC) residual_zeros<float>(h, L0, L0.rho32.p, s);
      else residual_zeros<bf16>(h, L0, L0.rho16.p, s);
    }
    if (prec == WM_PREC_FP64 && P.mgform == 1 && P.up >= 1 && repin_skippable(h, P)) {
      // slots of the two fused half sweeps: at most one per wavefront of
      // rows x z chunks of the dense part plus one per 64 records, with room
      const RBGeom &g0 = L0.g;
      long long nrec = 0;
      if (P.lcnd) nrec = cond_of<double>(L0).sor[0].n + cond_of<double>(L0).sor[1].n;
      long long cap = 2 * ((long long)(g0.W / 64 + 1) * (g0.PY + 16) * g0.PZ) + nrec / 64 + 4096;
      if ((long long)h->convparts.n < cap) {
        h->drop_graphs();  // captured graphs hold the old buffer's address
        h->gkey = wm_mg3d::GKey();
        h->convparts.alloc((size_t)cap);
      }
    }
    int iters = 0;
    double err = 2. * mgtol + 1.;
    static const bool eager_env = getenv("WARPMI_EAGER") != nullptr;
    if (mgmaxiters > 0 && (eager_env || h->prof.on)) {
      // Plain stream launches, one V-cycle at a time (and the event timing
      // of the finest-level half sweeps when it is switched on); a V-cycle
      // that exchanges the level-0 buffers does so as it is issued.
      const bool noskip = !repin_skippable(h, P);
      for (int k = 0; k < mgmaxiters; k++) {
        h->prof.used = 0;
        iteration(h, prec, P, mgtol, (int)mgmaxiters, s, (k == 0 && !reuse) || noskip);
        WM_HIP(hipMemcpyAsync(h->state_h, h->state.p, sizeof(SolveState), hipMemcpyDeviceToHost, s));
        WM_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < h->prof.used; i++) {
          float ms = 0.f;
          WM_HIP(hipEventElapsedTime(&ms, h->prof.ev[2 * i], h->prof.ev[2 * i + 1]));
          h->prof.ms += ms;
          h->prof.count++;
        }
        h->prof.used = 0;
        if (h->state_h->done) break;
      }
      iters = h->state_h->iters;
      err = h->state_h->err;
      h->last_iters = iters;
    } else if (mgmaxiters > 0) {
      // Capture one iteration as a graph (rebuilt when the parameters change).
      wm_mg3d::GKey key;
      key.prec = prec;
      key.mgform = (int)mgform;
      key.down = (int)downpasses;
      key.up = (int)uppasses;
      key.lcnd = (int)lcndbndy;
      key.icnd = (int)icndbndy;
      key.maxit = (int)mgmaxiters;
      key.mgparam = mgparam;
      key.mgtol = mgtol;
      const int o0 = h->orient;
      const bool pp = pp_now;
      if (!h->gexec[o0] || !(h->gkey == key)) {
        h->drop_graphs();
        // gexec: one V-cycle and the convergence test; gexec_rest: the
        // same without the first pin, for the V-cycles after the first
        // when repin_skippable holds.  Capturing a V-cycle that exchanges
        // the level-0 buffers exchanges them on the host too (the graph
        // holds the addresses of that orientation), so the graphs of both
        // orientations are captured, and the host's ends where it began.
        auto capture = [&](hipGraphExec_t *out, bool repin) {
          hipGraph_t graph;
          WM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          try {
            iteration(h, prec, P, mgtol, (int)mgmaxiters, s, repin);
          } catch (...) {
            hipGraph_t g2;
            (void)hipStreamEndCapture(s, &g2);
            if (g2) (void)hipGraphDestroy(g2);
            throw;
          }
          WM_HIP(hipStreamEndCapture(s, &graph));
          hipError_t e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          WM_HIP(e);
        };
        capture(&h->gexec[o0], true);
        if (prec == WM_PREC_FP64 && P.mgform == 1 && repin_skippable(h, P)) {
          capture(&h->gexec_rest[o0], false);
          if (pp) {
            WM_REQUIRE(h->orient != o0, "mg3d: the two-buffer V-cycle did not exchange");
            capture(&h->gexec[o0 ^ 1], true);
            capture(&h->gexec_rest[o0 ^ 1], false);
          }
        }
        WM_REQUIRE(h->orient == o0, "mg3d: graph capture changed the buffer orientation");
        h->gkey = key;
      }
      int batch = std::max(1, std::min<int>(h->last_iters > 0 ? h->last_iters : 4, (int)mgmaxiters));
      int launched = 0;
      int o = o0;  // orientation each launch starts from, as if every one runs
      // the first batch: its V-cycles of the gexec_rest form as one graph
      // (captured once per count and orientation: capturing exchanges the
      // buffers on the host once per V-cycle, so it ends where it began
      // after one more exchange when the count is odd)
      const int nmulti = h->gexec_rest[o0] ? batch - (reuse ? 0 : 1) : 0;
      hipGraphExec_t gm = nullptr;
      if (nmulti >= 2 && nmulti <= 16) {
        auto it = h->gmulti.find({o0, nmulti});
        if (it != h->gmulti.end()) {
          gm = it->second;
        } else {
          hipGraph_t graph;
          WM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          try {
            for (int k = 0; k < nmulti; k++)
              iteration(h, prec, P, mgtol, (int)mgmaxiters, s, false);
          } catch (...) {
            hipGraph_t g2;
            (void)hipStreamEndCapture(s, &g2);
            if (g2) (void)hipGraphDestroy(g2);
            throw;
          }
          WM_HIP(hipStreamEndCapture(s, &graph));
          hipError_t e = hipGraphInstantiate(&gm, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          WM_HIP(e);
          if (pp && (nmulti & 1)) pp_swap(h);
          WM_REQUIRE(h->orient == o0, "mg3d: graph capture changed the buffer orientation");
          h->gmulti[{o0, nmulti}] = gm;
        }
      }
      if (gm) {
        if (!reuse) {
          WM_HIP(hipGraphLaunch(h->gexec[o], s));
          launched++;
        }
        WM_HIP(hipGraphLaunch(gm, s));
        launched += nmulti;
        if (pp && (nmulti & 1)) o ^= 1;
        batch = 0;
      }
      for (;;) {
        for (int b = 0; b < batch && launched < mgmaxiters; b++, launched++) {
          const bool rest = (launched > 0 || reuse) && h->gexec_rest[o];
          WM_HIP(hipGraphLaunch(rest ? h->gexec_rest[o] : h->gexec[o], s));
          if (rest && pp) o ^= 1;
        }
        WM_HIP(hipMemcpyAsync(h->state_h, h->state.p, sizeof(SolveState), hipMemcpyDeviceToHost, s));
        WM_HIP(hipStreamSynchronize(s));
        if (h->state_h->done || launched >= mgmaxiters) break;
        batch = 1;
      }
      iters = h->state_h->iters;
      err = h->state_h->err;
      h->last_iters = iters;
      // the launches after the converged one changed nothing: the V-cycles
      // that ran, iters of them, exchanged the buffers once each but the
      // first (which pins and copies) unless the solve reused phi
      if (pp && h->gexec_rest[o0]) {
        const int swaps = std::max(0, iters - (reuse ? 0 : 1));
        if (swaps & 1) pp_swap(h);
      }
    }
    // Outputs: phi in natural layout, Dirichlet guards filled
    // (applyboundaryconditions3d with lwithdirichlet, f3d_mgrid.F:441-444).
    hipLaunchKernelGGL((k_rb_to_nat<double>), dim3(blocks(nn)), dim3(256), 0, s, L0.phi64.p, phi,
                       L0.g);
    WM_LAUNCHED();
    launch_bc<double>(phi, NatIdx{NatGeom{L0.nx, L0.ny, L0.nz}}, L0.nx, L0.ny, L0.nz, h->bounds,
                      1, 0, nullptr, s);
    WM_HIP(hipEventRecord(h->ev_out, s));
    WM_HIP(hipStreamWaitEvent(us, h->ev_out, 0));
    WM_HIP(hipStreamSynchronize(s));
    *mgiters = iters;
    *mgerror = err;
    h->last_ok = fp64plain && alldir && iters > 0 && repin_skippable(h, P);
    h->last_pp = pp_now;
    h->last_phi = phi;
  });
}

int wm_mg3d_reuse_phi(wm_mg3d *h) {
  WM_ENTRY({
    WM_REQUIRE(h != nullptr, "wm_mg3d_reuse_phi: NULL handle");
    h->reuse_next = true;
  });
}

int wm_mg3d_last_reused(const wm_mg3d *h, int64_t *reused) {
  WM_ENTRY({
    WM_REQUIRE(h && reused, "wm_mg3d_last_reused: NULL argument");
    *reused = h->last_reused ? 1 : 0;
  });
}

int wm_mg3d_tail_start(const wm_mg3d *h, int64_t *level) {
  WM_ENTRY({
    WM_REQUIRE(h && level, "wm_mg3d_tail_start: NULL argument");
    *level = h->tail_used ? h->tail.start : -1;
  });
}

int wm_phi_extrapolate(double *phi, double *prev, int64_t n, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(n >= 0, "wm_phi_extrapolate: n < 0");
    if (n == 0) return 0;
    WM_REQUIRE(phi && prev, "wm_phi_extrapolate: NULL array");
    hipLaunchKernelGGL(k_extrapolate, dim3((unsigned)cdiv(n, 256)), dim3(256), 0,
                       wm_stream(stream), phi, prev, (long long)n);
    WM_LAUNCHED();
  });
}

int wm_mg3d_set_sor_form(wm_mg3d *h, int64_t form) {
  WM_ENTRY({
    WM_REQUIRE(h != nullptr, "wm_mg3d_set_sor_form: NULL handle");
    WM_REQUIRE(form >= 0 && form <= 2, "wm_mg3d_set_sor_form: form must be 0, 1 or 2");
    if (h->sor_form != (int)form) {
      h->drop_graphs();  // captured V-cycles hold the other form's launches
      h->gkey = wm_mg3d::GKey();
      h->last_ok = false;
      h->sor_form = (int)form;
    }
  });
}

int wm_mg3d_set_timing(wm_mg3d *h, int64_t on) {
  WM_ENTRY({
    WM_REQUIRE(h != nullptr, "wm_mg3d_set_timing: NULL handle");
    h->prof.on = on != 0;
    h->prof.ms = 0.;
    h->prof.count = 0;
    h->prof.fused = false;
  });
}

int wm_mg3d_timing_fused(const wm_mg3d *h, int64_t *fused) {
  WM_ENTRY({
    WM_REQUIRE(h && fused, "wm_mg3d_timing_fused: NULL argument");
    *fused = h->prof.fused ? 1 : 0;
  });
}

int wm_mg3d_get_timing(const wm_mg3d *h, double *ms_total, int64_t *launches) {
  WM_ENTRY({
    WM_REQUIRE(h && ms_total && launches, "wm_mg3d_get_timing: NULL argument");
    *ms_total = h->prof.ms;
    *launches = h->prof.count;
  });
}

int wm_mg3d_sorhalfpass(wm_mg3d *h, int64_t level, int64_t parity, double *phi, const double *rho,
                        double mgparam, int64_t mgform, int64_t lcndbndy, int64_t icndbndy,
                        void *stream) {
  WM_ENTRY({
    WM_REQUIRE(h && phi && rho, "wm_mg3d_sorhalfpass: NULL argument");
    WM_REQUIRE(level >= 0 && level < (int64_t)h->lv.size(), "wm_mg3d_sorhalfpass: bad level");
    WM_REQUIRE(parity == 0 || parity == 1, "wm_mg3d_sorhalfpass: parity must be 0 or 1");
    // this call writes the level's red-black buffers (phi64 among them):
    // the next solve cannot start from the last one's state
    h->last_ok = false;
    hipStream_t s = wm_stream(stream);
    Level &L = *h->lv[level];
    Params P{mgparam, (int)mgform, 1, 1, (int)lcndbndy, (int)icndbndy};
    ensure_buffers<double>(L, L.coarser);
    build_mask(h, L, (int)level, P);
    ensure_records<double>(h, L, (int)level, P, false, 0);
    long long nn = (long long)(L.nx + 3) * (L.ny + 3) * (L.nz + 3);
    long long nr = (long long)(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
    hipLaunchKernelGGL((k_nat_to_rb<double>), dim3(blocks(nn)), dim3(256), 0, s, phi, L.phi64.p, L.g,
                       (double *)nullptr, L.phi64b.p);
    hipLaunchKernelGGL((k_rho_to_rb<double>), dim3(blocks(nr)), dim3(256), 0, s, rho, L.rho64.p, L.g);
    WM_LAUNCHED();
    halfpass<double>(h, L, (int)level, P, (int)parity, nullptr, s, false);
    hipLaunchKernelGGL((k_rb_to_nat<double>), dim3(blocks(nn)), dim3(256), 0, s, L.phi64.p, phi, L.g);
    WM_LAUNCHED();
  });
}

int wm_mg3d_sorpass(wm_mg3d *h, int64_t level, double *phi, const double *rho, double mgparam,
                    int64_t mgform, int64_t lcndbndy, int64_t icndbndy, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(h && phi && rho, "wm_mg3d_sorpass: NULL argument");
    WM_REQUIRE(level >= 0 && level < (int64_t)h->lv.size(), "wm_mg3d_sorpass: bad level");
    // this call writes the level's red-black buffers (phi64 among them):
    // the next solve cannot start from the last one's state
    h->last_ok = false;
    hipStream_t s = wm_stream(stream);
    Level &L = *h->lv[level];
    Params P{mgparam, (int)mgform, 1, 1, (int)lcndbndy, (int)icndbndy};
    ensure_buffers<double>(L, L.coarser);
    build_mask(h, L, (int)level, P);
    ensure_records<double>(h, L, (int)level, P, false, 0);
    long long nn = (long long)(L.nx + 3) * (L.ny + 3) * (L.nz + 3);
    long long nr = (long long)(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
    hipLaunchKernelGGL((k_nat_to_rb<double>), dim3(blocks(nn)), dim3(256), 0, s, phi, L.phi64.p, L.g,
                       (double *)nullptr, L.phi64b.p);
    hipLaunchKernelGGL((k_rho_to_rb<double>), dim3(blocks(nr)), dim3(256), 0, s, rho, L.rho64.p, L.g);
    WM_LAUNCHED();
    double *first = L.phi64.p;
    sorpass<double>(h, L, (int)level, P, nullptr, s);
    hipLaunchKernelGGL((k_rb_to_nat<double>), dim3(blocks(nn)), dim3(256), 0, s, L.phi64.p, phi, L.g);
    WM_LAUNCHED();
    // a fused pass left phi in the second buffer: restore the assignment
    // the solver's captured V-cycle graphs were recorded with (the next
    // solve converts its phi into both buffers anyway)
    if (L.phi64.p != first) L.phi64.swap(L.phi64b);
  });
}

int wm_mg3d_residual(wm_mg3d *h, int64_t level, const double *phi, const double *rho, double *res,
                     double mgparam, int64_t mgform, int64_t mgform2init, int64_t lcndbndy,
                     int64_t icndbndy, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(h && phi && rho && res, "wm_mg3d_residual: NULL argument");
    WM_REQUIRE(level >= 0 && level < (int64_t)h->lv.size(), "wm_mg3d_residual: bad level");
    // this call writes the level's red-black buffers (phi64 among them):
    // the next solve cannot start from the last one's state
    h->last_ok = false;
    hipStream_t s = wm_stream(stream);
    Level &L = *h->lv[level];
    Params P{mgparam, (int)mgform, 1, 1, (int)lcndbndy, (int)icndbndy};
    ensure_buffers<double>(L, L.coarser);
    build_mask(h, L, (int)level, P);
    long long nn = (long long)(L.nx + 3) * (L.ny + 3) * (L.nz + 3);
    long long nr = (long long)(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
    hipLaunchKernelGGL((k_nat_to_rb<double>), dim3(blocks(nn)), dim3(256), 0, s, phi, L.phi64.p, L.g,
                       (double *)nullptr, L.phi64b.p);
    hipLaunchKernelGGL((k_rho_to_rb<double>), dim3(blocks(nr)), dim3(256), 0, s, rho, L.rho64.p, L.g);
    WM_LAUNCHED();
    L.res64.alloc_rb(std::max<size_t>(2 * L.g.cs, (size_t)nn));
    residual_rb<double, double>(h, L, (int)level, P, L.phi64.p, L.rho64.p, L.res64.p,
                                mgform2init ? 1 : 0, nullptr, s);
    hipLaunchKernelGGL((k_rb_to_nat<double>), dim3(blocks(nn)), dim3(256), 0, s, L.res64.p, res,
                       L.g);
    WM_LAUNCHED();
  });
}

int wm_mg3d_restrict(wm_mg3d *h, int64_t level, const double *res, double *rhocoarse,
                     void *stream) {
  WM_ENTRY({
    WM_REQUIRE(h && res && rhocoarse, "wm_mg3d_restrict: NULL argument");
    WM_REQUIRE(level >= 0 && level + 1 < (int64_t)h->lv.size(), "wm_mg3d_restrict: no coarser level");
    Level &L = *h->lv[level];
    Level &Lc = *h->lv[level + 1];
    restrict_to<double, double>(L, res, NatIdx{NatGeom{L.nx, L.ny, L.nz}}, rhocoarse,
                                RhoIdx{Lc.nx, Lc.ny}, nullptr, wm_stream(stream), Lc.nx, Lc.ny,
                                Lc.nz);
  });
}

int wm_mg3d_expand(wm_mg3d *h, int64_t level, double *phi, const double *phicoarse, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(h && phi && phicoarse, "wm_mg3d_expand: NULL argument");
    WM_REQUIRE(level >= 0 && level + 1 < (int64_t)h->lv.size(), "wm_mg3d_expand: no coarser level");
    Level &L = *h->lv[level];
    Level &Lc = *h->lv[level + 1];
    expand_from<double>(L, phi, NatIdx{NatGeom{L.nx, L.ny, L.nz}}, phicoarse,
                        NatIdx{NatGeom{Lc.nx, Lc.ny, Lc.nz}}, nullptr, wm_stream(stream));
  });
}

int wm_applyboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz, double *u,
                                 const int64_t *bounds, int64_t lwithdirichlet,
                                 int64_t lzerodirichlet, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(u && bounds, "wm_applyboundaryconditions3d: NULL argument");
    int b[6];
    for (int i = 0; i < 6; i++) b[i] = (int)bounds[i];
    launch_bc<double>(u, NatIdx{NatGeom{(int)nx, (int)ny, (int)nz}}, (int)nx, (int)ny, (int)nz, b,
                      (int)lwithdirichlet, (int)lzerodirichlet, nullptr, wm_stream(stream));
  });
}

}  // extern "C"
