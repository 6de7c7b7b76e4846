// Conductor-corrected field gather (f3d.lcorrectede): the device side of
// sete3dwithconductor (pywarp90/f3d_mgrid.F:4774-5010, GETEEFROME
// 5164-5233), shared by the standalone gather and the fused particle
// advance.  FP64 in the reference's operation order.
#pragma once
#include "wm_internal.h"

// icgrid value of an interior conductor node: above every record index
// (the reference uses max(nmax)+1, f3d_mgrid.F:4607-4614).
#define WM_CF_INTERIOR 2147483647

// Device view of a wm_condfield: the node classification icgrid
// (0:nx, 0:ny, 0:nz), x fastest: 0 vacuum, WM_CF_INTERIOR interior,
// +(r+1) even record r, -(r+1) odd record r; and per record the
// one-sided E, efield0 (3 per record), and its extrapolation into the
// conductor, efieldd (6 per record, mx,px,my,py,mz,pz).
struct CondView {
  const int *icgrid;
  const double *e0[2];
  const double *ed[2];
  int nx, ny, nz;
};

// What a gather needs besides the particle: the field data, phi
// (-1:n+1)^3, the grid E selfe (3, 0:n)^3 for efetch 3, and the mesh.
struct CondGather {
  CondView v;
  const double *phi, *selfe;
  int efetch3;
  double xmmin, ymmin, zmmin, zgrid;
  double dxi, dyi, dzi, tdxi, tdyi, tdzi;
};

// icgrid at (i, j, k); a node past the upper faces (reached with weight
// zero by a particle on that face) reads as vacuum.
__device__ inline int cf_icg(const CondView &c, long long i, long long j, long long k) {
  if (i > c.nx || j > c.ny || k > c.nz) return 0;
  return c.icgrid[i + (long long)(c.nx + 1) * (j + (long long)(c.ny + 1) * k)];
}

__device__ inline long long cf_clamp(long long v, long long lo, long long hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// phi(i, j, k) with the index kept inside (-1:n+1) (only weight-zero terms
// of a particle on an upper face reach past it).
__device__ inline double cf_phi(const CondGather &g, long long i, long long j, long long k) {
  const CondView &c = g.v;
  i = cf_clamp(i, -1, c.nx + 1);
  j = cf_clamp(j, -1, c.ny + 1);
  k = cf_clamp(k, -1, c.nz + 1);
  return g.phi[(i + 1) + (long long)(c.nx + 3) * ((j + 1) + (long long)(c.ny + 3) * (k + 1))];
}

// selfe(comp, i, j, k), index kept inside (0:n).
__device__ inline double cf_se(const CondGather &g, int comp, long long i, long long j,
                               long long k) {
  const CondView &c = g.v;
  i = cf_clamp(i, 0, c.nx);
  j = cf_clamp(j, 0, c.ny);
  k = cf_clamp(k, 0, c.nz);
  return g.selfe[3 * (i + (long long)(c.nx + 1) * (j + (long long)(c.ny + 1) * k)) + comp];
}

__device__ inline double cf_e0(const CondView &c, int ic, int id) {
  return ic < 0 ? c.e0[1][3 * (long long)(-ic - 1) + id] : c.e0[0][3 * (long long)(ic - 1) + id];
}

__device__ inline double cf_ed(const CondView &c, int ic, int k) {
  return ic < 0 ? c.ed[1][6 * (long long)(-ic - 1) + k] : c.ed[0][6 * (long long)(ic - 1) + k];
}

// GETEEFROME: the contribution of the node pair (i+o, i+o+p) to component
// id with weights wm, wp (f3d_mgrid.F:5164-5214).
__device__ inline void cf_getee(const CondGather &g, long long i, long long j, long long k,
                                int id, int ox, int oy, int oz, int px, int py, int pz,
                                double di, double wm, double wp, double &ee, double &ww) {
  const CondView &c = g.v;
  long long a0 = i + ox, b0 = j + oy, c0k = k + oz;
  long long a1 = a0 + px, b1 = b0 + py, c1k = c0k + pz;
  int c0 = cf_icg(c, a0, b0, c0k), c1 = cf_icg(c, a1, b1, c1k);
  if (c0 == 0) {
    if (g.efetch3)
      ee = ee + wm * cf_se(g, id, a0, b0, c0k);
    else
      ee = ee + wm * (cf_phi(g, a0 - px, b0 - py, c0k - pz) - cf_phi(g, a1, b1, c1k)) * 0.5 * di;
    ww = ww + wm;
  } else if (c0 != WM_CF_INTERIOR) {
    ee = ee + wm * cf_e0(c, c0, id);
    ww = ww + wm;
    if (c1 == WM_CF_INTERIOR) {
      ee = ee + wp * cf_ed(c, c0, 2 * id + 1);
      ww = ww + wp;
    }
  }
  if (c1 == 0) {
    if (g.efetch3)
      ee = ee + wp * cf_se(g, id, a1, b1, c1k);
    else
      ee = ee + wp * (cf_phi(g, a0, b0, c0k) - cf_phi(g, a1 + px, b1 + py, c1k + pz)) * 0.5 * di;
    ww = ww + wp;
  } else if (c1 != WM_CF_INTERIOR) {
    ee = ee + wp * cf_e0(c, c1, id);
    ww = ww + wp;
    if (c0 == WM_CF_INTERIOR) {
      ee = ee + wm * cf_ed(c, c1, 2 * id);
      ww = ww + wm;
    }
  }
}

// E at (x, y, z) as sete3dwithconductor computes it for efetch 1 or 3.
// Returns false for a position outside [0, n] on some axis (the reference
// skips the particle); NaN positions are rejected too.
__device__ inline bool gather_e_cond(const CondGather &g, double x, double y, double z,
                                     double &ext, double &eyt, double &ezt) {
  const CondView &c = g.v;
  double xi = (x - g.xmmin) * g.dxi;
  double yj = (y - g.ymmin) * g.dyi;
  double zk = (z - g.zgrid - g.zmmin) * g.dzi;
  if (!(xi >= 0. && xi <= (double)c.nx && yj >= 0. && yj <= (double)c.ny && zk >= 0. &&
        zk <= (double)c.nz))
    return false;
  long long i = (long long)xi, j = (long long)yj, k = (long long)zk;
  double u1 = xi - i, v1 = yj - j, w1 = zk - k;
  double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
  bool vac = cf_icg(c, i, j, k) == 0 && cf_icg(c, i + 1, j, k) == 0 &&
             cf_icg(c, i, j + 1, k) == 0 && cf_icg(c, i + 1, j + 1, k) == 0 &&
             cf_icg(c, i, j, k + 1) == 0 && cf_icg(c, i + 1, j, k + 1) == 0 &&
             cf_icg(c, i, j + 1, k + 1) == 0 && cf_icg(c, i + 1, j + 1, k + 1) == 0;
  if (vac) {
    if (g.efetch3) {
      double out[3];
#pragma unroll
      for (int q = 0; q < 3; q++)
        out[q] = u0 * v0 * w0 * cf_se(g, q, i, j, k) + u1 * v0 * w0 * cf_se(g, q, i + 1, j, k) +
                 u0 * v1 * w0 * cf_se(g, q, i, j + 1, k) +
                 u1 * v1 * w0 * cf_se(g, q, i + 1, j + 1, k) +
                 u0 * v0 * w1 * cf_se(g, q, i, j, k + 1) +
                 u1 * v0 * w1 * cf_se(g, q, i + 1, j, k + 1) +
                 u0 * v1 * w1 * cf_se(g, q, i, j + 1, k + 1) +
                 u1 * v1 * w1 * cf_se(g, q, i + 1, j + 1, k + 1);
      ext = out[0];
      eyt = out[1];
      ezt = out[2];
    } else {
      double tx = g.tdxi, ty = g.tdyi, tz = g.tdzi;
#define PH(a, b, cc) cf_phi(g, i + (a), j + (b), k + (cc))
      ext = u0 * v0 * w0 * (PH(-1, 0, 0) - PH(1, 0, 0)) * tx +
            u1 * v0 * w0 * (PH(0, 0, 0) - PH(2, 0, 0)) * tx +
            u0 * v1 * w0 * (PH(-1, 1, 0) - PH(1, 1, 0)) * tx +
            u1 * v1 * w0 * (PH(0, 1, 0) - PH(2, 1, 0)) * tx +
            u0 * v0 * w1 * (PH(-1, 0, 1) - PH(1, 0, 1)) * tx +
            u1 * v0 * w1 * (PH(0, 0, 1) - PH(2, 0, 1)) * tx +
            u0 * v1 * w1 * (PH(-1, 1, 1) - PH(1, 1, 1)) * tx +
            u1 * v1 * w1 * (PH(0, 1, 1) - PH(2, 1, 1)) * tx;
      eyt = u0 * v0 * w0 * (PH(0, -1, 0) - PH(0, 1, 0)) * ty +
            u1 * v0 * w0 * (PH(1, -1, 0) - PH(1, 1, 0)) * ty +
            u0 * v1 * w0 * (PH(0, 0, 0) - PH(0, 2, 0)) * ty +
            u1 * v1 * w0 * (PH(1, 0, 0) - PH(1, 2, 0)) * ty +
            u0 * v0 * w1 * (PH(0, -1, 1) - PH(0, 1, 1)) * ty +
            u1 * v0 * w1 * (PH(1, -1, 1) - PH(1, 1, 1)) * ty +
            u0 * v1 * w1 * (PH(0, 0, 1) - PH(0, 2, 1)) * ty +
            u1 * v1 * w1 * (PH(1, 0, 1) - PH(1, 2, 1)) * ty;
      ezt = u0 * v0 * w0 * (PH(0, 0, -1) - PH(0, 0, 1)) * tz +
            u1 * v0 * w0 * (PH(1, 0, -1) - PH(1, 0, 1)) * tz +
            u0 * v1 * w0 * (PH(0, 1, -1) - PH(0, 1, 1)) * tz +
            u1 * v1 * w0 * (PH(1, 1, -1) - PH(1, 1, 1)) * tz +
            u0 * v0 * w1 * (PH(0, 0, 0) - PH(0, 0, 2)) * tz +
            u1 * v0 * w1 * (PH(1, 0, 0) - PH(1, 0, 2)) * tz +
            u0 * v1 * w1 * (PH(0, 1, 0) - PH(0, 1, 2)) * tz +
            u1 * v1 * w1 * (PH(1, 1, 0) - PH(1, 1, 2)) * tz;
#undef PH
    }
    return true;
  }
  double ex = 0., ey = 0., ez = 0., wx = 0., wy = 0., wz = 0.;
  cf_getee(g, i, j, k, 0, 0, 0, 0, 1, 0, 0, g.dxi, u0 * v0 * w0, u1 * v0 * w0, ex, wx);
  cf_getee(g, i, j, k, 0, 0, 1, 0, 1, 0, 0, g.dxi, u0 * v1 * w0, u1 * v1 * w0, ex, wx);
  cf_getee(g, i, j, k, 0, 0, 0, 1, 1, 0, 0, g.dxi, u0 * v0 * w1, u1 * v0 * w1, ex, wx);
  cf_getee(g, i, j, k, 0, 0, 1, 1, 1, 0, 0, g.dxi, u0 * v1 * w1, u1 * v1 * w1, ex, wx);
  if (wx != 0.) wx = 1. / wx;
  cf_getee(g, i, j, k, 1, 0, 0, 0, 0, 1, 0, g.dyi, u0 * v0 * w0, u0 * v1 * w0, ey, wy);
  cf_getee(g, i, j, k, 1, 1, 0, 0, 0, 1, 0, g.dyi, u1 * v0 * w0, u1 * v1 * w0, ey, wy);
  cf_getee(g, i, j, k, 1, 0, 0, 1, 0, 1, 0, g.dyi, u0 * v0 * w1, u0 * v1 * w1, ey, wy);
  cf_getee(g, i, j, k, 1, 1, 0, 1, 0, 1, 0, g.dyi, u1 * v0 * w1, u1 * v1 * w1, ey, wy);
  if (wy != 0.) wy = 1. / wy;
  cf_getee(g, i, j, k, 2, 0, 0, 0, 0, 0, 1, g.dzi, u0 * v0 * w0, u0 * v0 * w1, ez, wz);
  cf_getee(g, i, j, k, 2, 1, 0, 0, 0, 0, 1, g.dzi, u1 * v0 * w0, u1 * v0 * w1, ez, wz);
  cf_getee(g, i, j, k, 2, 0, 1, 0, 0, 0, 1, g.dzi, u0 * v1 * w0, u0 * v1 * w1, ez, wz);
  cf_getee(g, i, j, k, 2, 1, 1, 0, 0, 0, 1, g.dzi, u1 * v1 * w0, u1 * v1 * w1, ez, wz);
  if (wz != 0.) wz = 1. / wz;
  ext = ex * wx;
  eyt = ey * wy;
  ezt = ez * wz;
  return true;
}

struct wm_condfield;
// The device view of a handle (defined with the handle).
CondView wm_condfield_view(const wm_condfield *cf);
// The gather arguments for a handle; checks efetch (1 or 3) and that the
// array it reads (phi or selfe) is given.
CondGather wm_condfield_gather(const wm_condfield *cf, const double *phi, const double *selfe,
                               int64_t efetch, double zgrid, double xmmin, double ymmin,
                               double zmmin, double dx, double dy, double dz);
