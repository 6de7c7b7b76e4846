// Conductor-corrected E field for gfx950 (f3d.lcorrectede): the node
// classification of setupconductorfielddata, the one-sided E at the
// subgrid points of getefieldatconductorsubgrid, fixefieldatconductorpoints
// and the gather of sete3dwithconductor (pywarp90/f3d_mgrid.F:4581-5233),
// for the finest level (the particles' grid), efetch 1 and 3, no
// symmetries.  FP64 in the reference's operation order.
#include "condfield.h"
#include "../../include/warpmi.h"

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <vector>

struct wm_condfield {
  int nx = 0, ny = 0, nz = 0;
  DevBuf<int> icgrid;
  int nsub[2] = {0, 0};
  DevBuf<int> pos[2];        // record node (ix, iy, iz)
  DevBuf<double> dels[2];    // 6 per record, mx,px,my,py,mz,pz
  DevBuf<double> volt[2];
  DevBuf<double> e0[2], ed[2];
  bool skip_efield = false;  // one parity list is empty (all levels)
  int nfix = 0;
  DevBuf<long long> fixnode;  // distinct interior nodes of the finest level
  DevBuf<double> fixfac;      // 2 to the power of the node's multiplicity
};

CondView wm_condfield_view(const wm_condfield *cf) {
  CondView v;
  v.icgrid = cf->icgrid.p;
  for (int p = 0; p < 2; p++) {
    v.e0[p] = cf->e0[p].p;
    v.ed[p] = cf->ed[p].p;
  }
  v.nx = cf->nx;
  v.ny = cf->ny;
  v.nz = cf->nz;
  return v;
}

namespace {

struct RecSet {
  const int *pos;
  const double *dels, *volt;
  double *e0, *ed;
  int n;
};

__device__ inline double cfe_phi(const double *phi, int nx, int ny, long long i, long long j,
                                 long long k) {
  return phi[(i + 1) + (long long)(nx + 3) * ((j + 1) + (long long)(ny + 3) * (k + 1))];
}

// getefieldatconductorsubgrid / SETEE (f3d_mgrid.F:4635-4735): one thread
// per record of both lists (even first).  efieldd is written only where
// the neighbour on that side is an interior node; the other entries keep
// what they held.
__global__ __launch_bounds__(256) void k_cf_efield(const double *__restrict__ phi,
                                                   const int *__restrict__ icgrid, int nx, int ny,
                                                   int nz, RecSet r0, RecSet r1, double dxi,
                                                   double dyi, double dzi) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)r0.n + r1.n) return;
  const RecSet &r = t < r0.n ? r0 : r1;
  long long ic = t < r0.n ? t : t - r0.n;
  int ix = r.pos[3 * ic], iy = r.pos[3 * ic + 1], iz = r.pos[3 * ic + 2];
  const double *d = r.dels + 6 * ic, *v = r.volt + 6 * ic;
  auto icg = [&](long long a, long long b, long long c) {
    return icgrid[a + (long long)(nx + 1) * (b + (long long)(ny + 1) * c)];
  };
  double ppp = cfe_phi(phi, nx, ny, ix, iy, iz);
#pragma unroll
  for (int id = 0; id < 3; id++) {
    int mx = id == 0 ? -1 : 0, my = id == 1 ? -1 : 0, mz = id == 2 ? -1 : 0;
    double di = id == 0 ? dxi : (id == 1 ? dyi : dzi);
    int im = 2 * id, ip = 2 * id + 1;
    double vv0, dd0, vv1, dd1;
    if (0. < d[im] && d[im] <= 1.) {
      vv0 = v[im];
      dd0 = d[im];
    } else {
      vv0 = cfe_phi(phi, nx, ny, ix + mx, iy + my, iz + mz);
      dd0 = 1.;
    }
    if (0. < d[ip] && d[ip] <= 1.) {
      vv1 = v[ip];
      dd1 = d[ip];
    } else {
      vv1 = cfe_phi(phi, nx, ny, ix - mx, iy - my, iz - mz);
      dd1 = 1.;
    }
    r.e0[3 * ic + id] =
        (dd1 / (dd0 + dd1) * (vv0 - ppp) / dd0 + dd0 / (dd0 + dd1) * (ppp - vv1) / dd1) * di;
    if (ix + mx >= 0 && iy + my >= 0 && iz + mz >= 0 &&
        icg(ix + mx, iy + my, iz + mz) == WM_CF_INTERIOR)
      r.ed[6 * ic + im] = ((vv0 - ppp) / dd0 - (2. - dd0) / (2. + dd1) * (ppp - vv1) / dd1) *
                          ((2. + dd1) / (dd1 + dd0)) * di;
    if (ix - mx <= nx && iy - my <= ny && iz - mz <= nz &&
        icg(ix - mx, iy - my, iz - mz) == WM_CF_INTERIOR)
      r.ed[6 * ic + ip] = ((ppp - vv1) / dd1 - (2. - dd1) / (2. + dd0) * (vv0 - ppp) / dd0) *
                          ((2. + dd0) / (dd0 + dd1)) * di;
  }
}

// fixefieldatconductorpoints (f3d_mgrid.F:4737-4772): the grid E doubled
// at each interior node, once per listing (one thread per distinct node,
// scaled by 2^multiplicity, which is the same exact power-of-two scaling).
__global__ __launch_bounds__(256) void k_cf_fix(double *__restrict__ selfe, int n,
                                                const long long *__restrict__ node,
                                                const double *__restrict__ fac) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  double *e = selfe + 3 * node[t];
  double f = fac[t];
  e[0] = f * e[0];
  e[1] = f * e[1];
  e[2] = f * e[2];
}

__global__ __launch_bounds__(256) void k_cf_sete(CondGather g, long long np,
                                                 const double *__restrict__ xp,
                                                 const double *__restrict__ yp,
                                                 const double *__restrict__ zp,
                                                 double *__restrict__ ex, double *__restrict__ ey,
                                                 double *__restrict__ ez) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  double a, b, c;
  if (!gather_e_cond(g, xp[ip], yp[ip], zp[ip], a, b, c)) return;
  ex[ip] = ex[ip] + a;
  ey[ip] = ey[ip] + b;
  ez[ip] = ez[ip] + c;
}

}  // namespace

CondGather wm_condfield_gather(const wm_condfield *cf, const double *phi, const double *selfe,
                               int64_t efetch, double zgrid, double xmmin, double ymmin,
                               double zmmin, double dx, double dy, double dz) {
  WM_REQUIRE(cf != nullptr, "NULL wm_condfield");
  WM_REQUIRE(efetch == 1 || efetch == 3, "conductor-corrected gather: efetch must be 1 or 3");
  WM_REQUIRE(efetch == 1 ? phi != nullptr : selfe != nullptr,
             "conductor-corrected gather: efetch 1 needs phi, efetch 3 needs selfe");
  CondGather g;
  g.v = wm_condfield_view(cf);
  g.phi = phi;
  g.selfe = selfe;
  g.efetch3 = efetch == 3;
  g.xmmin = xmmin;
  g.ymmin = ymmin;
  g.zmmin = zmmin;
  g.zgrid = zgrid;
  g.dxi = 1. / dx;
  g.dyi = 1. / dy;
  g.dzi = 1. / dz;
  g.tdxi = 1. / (2. * dx);
  g.tdyi = 1. / (2. * dy);
  g.tdzi = 1. / (2. * dz);
  return g;
}

extern "C" {

int wm_condfield_create(wm_condfield **out, int64_t nx, int64_t ny, int64_t nz,
                        const wm_conductor_lists *c) {
  WM_ENTRY({
    WM_REQUIRE(out != nullptr, "wm_condfield_create: NULL out");
    *out = nullptr;
    WM_REQUIRE(nx > 0 && ny > 0 && nz > 0, "wm_condfield_create: grid must have n > 0");
    WM_REQUIRE((nx + 1) * (ny + 1) * (nz + 1) < (1ll << 31) / 3,
               "wm_condfield_create: grid too large");
    std::unique_ptr<wm_condfield> h(new wm_condfield);
    h->nx = (int)nx;
    h->ny = (int)ny;
    h->nz = (int)nz;
    size_t nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    std::vector<int> icg(nn, 0);
    auto node = [&](const int64_t *ix3, const char *what) -> long long {
      int64_t a = ix3[0], b = ix3[1], d = ix3[2];
      WM_REQUIRE(a >= 0 && a <= nx && b >= 0 && b <= ny && d >= 0 && d <= nz,
                 std::string("wm_condfield_create: ") + what + " point outside the grid");
      return a + (nx + 1) * (b + (ny + 1) * d);
    };
    std::map<long long, int> mult;
    if (c && c->int_n > 0) {
      WM_REQUIRE(c->int_indx && c->int_istart, "wm_condfield_create: NULL interior list");
      for (int64_t i = c->int_istart[0]; i < c->int_istart[1]; i++) {
        long long e = node(c->int_indx + 3 * i, "interior");
        icg[e] = WM_CF_INTERIOR;
        mult[e] += 1;
      }
    }
    for (int p = 0; p < 2; p++) {
      if (!c || c->sub_n[p] == 0) continue;
      WM_REQUIRE(c->sub_indx[p] && c->sub_dels[p] && c->sub_volt[p] && c->sub_istart[p],
                 "wm_condfield_create: NULL subgrid list");
      int64_t a = c->sub_istart[p][0], b = c->sub_istart[p][1];
      WM_REQUIRE(b - a < (1ll << 30), "wm_condfield_create: too many subgrid points");
      std::vector<int> pos;
      std::vector<double> dl, vt;
      for (int64_t i = a; i < b; i++) {
        long long e = node(c->sub_indx[p] + 3 * i, "subgrid");
        int r = (int)(i - a) + 1;
        icg[e] = p == 0 ? r : -r;
        for (int q = 0; q < 3; q++) pos.push_back((int)c->sub_indx[p][3 * i + q]);
        for (int q = 0; q < 6; q++) {
          dl.push_back(c->sub_dels[p][6 * i + q]);
          vt.push_back(c->sub_volt[p][6 * i + q]);
        }
      }
      h->nsub[p] = (int)(b - a);
      h->pos[p].upload(pos);
      h->dels[p].upload(dl);
      h->volt[p].upload(vt);
      h->e0[p].alloc((size_t)3 * h->nsub[p]);
      h->ed[p].alloc((size_t)6 * h->nsub[p]);
    }
    h->skip_efield = !c || c->sub_n[0] == 0 || c->sub_n[1] == 0;
    h->icgrid.upload(icg);
    std::vector<long long> fn;
    std::vector<double> ff;
    for (auto &kv : mult) {
      fn.push_back(kv.first);
      double f = 1.;
      for (int q = 0; q < kv.second; q++) f = 2. * f;
      ff.push_back(f);
    }
    h->nfix = (int)fn.size();
    h->fixnode.upload(fn);
    h->fixfac.upload(ff);
    *out = h.release();
  });
}

int wm_condfield_destroy(wm_condfield *cf) {
  WM_ENTRY({ delete cf; });
}

int wm_condfield_efield(wm_condfield *cf, const double *phi, double dx, double dy, double dz,
                        void *stream) {
  WM_ENTRY({
    WM_REQUIRE(cf && phi, "wm_condfield_efield: NULL argument");
    if (cf->skip_efield) return 0;
    int n = cf->nsub[0] + cf->nsub[1];
    if (n == 0) return 0;
    RecSet r[2];
    for (int p = 0; p < 2; p++)
      r[p] = RecSet{cf->pos[p].p, cf->dels[p].p, cf->volt[p].p, cf->e0[p].p, cf->ed[p].p,
                    cf->nsub[p]};
    hipLaunchKernelGGL(k_cf_efield, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, wm_stream(stream),
                       phi, cf->icgrid.p, cf->nx, cf->ny, cf->nz, r[0], r[1], 1. / dx, 1. / dy,
                       1. / dz);
    WM_LAUNCHED();
  });
}

int wm_condfield_fixefield(const wm_condfield *cf, double *selfe, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(cf && selfe, "wm_condfield_fixefield: NULL argument");
    if (cf->nfix == 0) return 0;
    hipLaunchKernelGGL(k_cf_fix, dim3((unsigned)cdiv(cf->nfix, 256)), dim3(256), 0,
                       wm_stream(stream), selfe, cf->nfix, cf->fixnode.p, cf->fixfac.p);
    WM_LAUNCHED();
  });
}

int wm_condfield_read(const wm_condfield *cf, int64_t *nsub, int32_t *icgrid, double *e0_even,
                      double *ed_even, double *e0_odd, double *ed_odd, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(cf && nsub, "wm_condfield_read: NULL argument");
    nsub[0] = cf->nsub[0];
    nsub[1] = cf->nsub[1];
    hipStream_t s = wm_stream(stream);
    auto copy = [&](void *dst, const void *src, size_t bytes) {
      if (dst && bytes) WM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
    };
    copy(icgrid, cf->icgrid.p, sizeof(int) * cf->icgrid.n);
    copy(e0_even, cf->e0[0].p, sizeof(double) * cf->e0[0].n);
    copy(ed_even, cf->ed[0].p, sizeof(double) * cf->ed[0].n);
    copy(e0_odd, cf->e0[1].p, sizeof(double) * cf->e0[1].n);
    copy(ed_odd, cf->ed[1].p, sizeof(double) * cf->ed[1].n);
    WM_HIP(hipStreamSynchronize(s));
  });
}

int wm_sete3d_conductor(const wm_condfield *cf, const double *phi, const double *selfe,
                        int64_t efetch, int64_t np, const double *xp, const double *yp,
                        const double *zp, double zgrid, double xmmin, double ymmin, double zmmin,
                        double dx, double dy, double dz, double *ex, double *ey, double *ez,
                        void *stream) {
  WM_ENTRY({
    CondGather g =
        wm_condfield_gather(cf, phi, selfe, efetch, zgrid, xmmin, ymmin, zmmin, dx, dy, dz);
    WM_REQUIRE(np >= 0, "wm_sete3d_conductor: np < 0");
    if (np == 0) return 0;
    WM_REQUIRE(xp && yp && zp && ex && ey && ez, "wm_sete3d_conductor: NULL array");
    hipLaunchKernelGGL(k_cf_sete, dim3((unsigned)cdiv(np, 256)), dim3(256), 0, wm_stream(stream),
                       g, (long long)np, xp, yp, zp, ex, ey, ez);
    WM_LAUNCHED();
  });
}

}  // extern "C"
