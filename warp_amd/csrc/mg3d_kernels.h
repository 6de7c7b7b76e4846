// Device kernels of the multigrid solver (pywarp90/f3d_mgrid.F), for
// gfx950.  Included by mg3d.hip only.
//
// Layout.  Each level's phi and rho live in a red-black split layout: the
// padded (-1:nx+1,-1:ny+1,-1:nz+1) array is split by the colour
// c = (ix+iy+iz) mod 2 into two arrays of rows of W = (nx+4)/2 elements.
// With px = ix+1 (likewise py, pz) the element of a point is
//   c*cs + (px>>1) + W*(py + PY*pz).
// A half sweep of colour c then streams one array (read+write) and reads
// the other: every cache line moved carries only useful data, so the
// HBM traffic of a half sweep is the algorithmic 2 scalars per node.
// Residuals stay in the reference's natural layout with one guard cell.
//
// Every kernel takes `skip`: when *skip != 0 the kernel returns at once.
// The solver sets it when the convergence test is met so that V-cycles
// enqueued ahead of the test become no-ops.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct RBGeom {
  int nx, ny, nz;  // cells of the level
  int W, PY, PZ;   // elements per half row, padded rows and planes
  long long cs;    // elements per colour array
};

__host__ __device__ inline long long rb_index(const RBGeom &g, int ix, int iy,
                                              int iz) {
  int px = ix + 1, py = iy + 1, pz = iz + 1;
  int c = (px + py + pz + 1) & 1;
  return c * g.cs + (px >> 1) + (long long)g.W * (py + (long long)g.PY * pz);
}

struct NatGeom {
  int nx, ny, nz;
};

__host__ __device__ inline long long nat_index(const NatGeom &g, int ix, int iy,
                                               int iz) {
  return (long long)(ix + 1) +
         (long long)(g.nx + 3) * ((long long)(iy + 1) + (long long)(g.ny + 3) * (iz + 1));
}

// Accessors so that the same kernel body serves both layouts.
// Element of point (x, y, z) (each in -1..n+1) of a red-black array.
__device__ inline int rb_el(const RBGeom &g, int x, int y, int z) {
  return ((x + y + z) & 1) * (int)g.cs + ((x + 1) >> 1) + g.W * ((y + 1) + g.PY * (z + 1));
}

struct RBIdx {
  RBGeom g;
  __device__ long long operator()(int ix, int iy, int iz) const {
    return rb_index(g, ix, iy, iz);
  }
};
struct NatIdx {
  NatGeom g;
  __device__ long long operator()(int ix, int iy, int iz) const {
    return nat_index(g, ix, iy, iz);
  }
};

#define WM_SKIP() \
  if (skip && *skip) return

// Storage types.  Level arrays hold S; arithmetic is done in St<S>::T.
// For double and float both are the same type and the conversions are the
// identity.  bf16 is the upper half of an IEEE single, kept as raw bits and
// rounded to nearest even on store.
struct bf16 {
  uint16_t u;
};
template <typename S>
struct St {
  typedef S T;
  __host__ __device__ static S ld(S v) { return v; }
  __host__ __device__ static S st(S v) { return v; }
};
template <>
struct St<bf16> {
  typedef float T;
  __device__ static float ld(bf16 v) { return __uint_as_float((uint32_t)v.u << 16); }
  __device__ static bf16 st(float f) {
    uint32_t x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return bf16{(uint16_t)((x >> 16) | 0x40u)};
    x += 0x7fffu + ((x >> 16) & 1u);
    return bf16{(uint16_t)(x >> 16)};
  }
};

// Side encodings of the subgrid-conductor stencil, one per neighbour
// (order mx,px,my,py,mz,pz), 2 bits each: the neighbour value is used raw,
// multiplied by k, divided by k, or replaced by the constant k.  This is
// the set of forms handlesubgrid2 (f3d_mgrid.F:2628-2764) and condbndymg
// icndbndy=1 (2846-2895) produce; the constants are computed on the host
// with the reference's expressions.
enum { SIDE_RAW = 0, SIDE_MUL = 1, SIDE_DIV = 2, SIDE_CONST = 3 };

template <typename T>
__device__ inline T side_value(uint32_t mode, T k, T raw) {
  if (mode == SIDE_RAW) return raw;
  if (mode == SIDE_MUL) return raw * k;
  if (mode == SIDE_DIV) return raw / k;
  return k;
}

// One colour's active subgrid points on one level (SoA, device).
template <typename T>
struct SubList {
  int n;
  const int *pos;        // (n,3) natural ix,iy,iz
  const uint32_t *modes; // 2 bits per side
  const T *k;            // (n,6)
  const T *denom;        // (n)
  const T *fac;          // (n) residual factor ppp (residual lists only)
  const uint32_t *zero;  // (n) 6-bit mask of neighbours whose residual is zeroed
};

// Largest of the wavefront's values (lane 0 holds it); NaN values are
// passed over, as by the convergence test's comparisons.
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// XCD-aware block order: hardware hands consecutive workgroups to the 8
// XCDs round robin, so block b runs on XCD b%8.  Remap so that each XCD
// walks a contiguous range of tiles and its L2 holds the neighbouring
// planes it needs.
__device__ inline int xcd_remap(int b, int nb) {
  int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// ---------------------------------------------------------------------------
// SOR half sweep (sorhalfpass3d, f3d_mgrid.F:2229-2243) fused with the
// subgrid-conductor update of the same colour (condbndymg 2766-2980).
// Dense part: one thread per colour-c point, 64 x 4 tiles.  Points in the
// conductor mask are left to the sparse part, which reads the point's value
// before this sweep (the reference's prevphi) and other-colour neighbours,
// neither of which the dense part writes.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_sor_halfpass(
    T *__restrict__ phi, const T *__restrict__ rho, RBGeom g, int c, int ixmin,
    int ixmax, int iymin, int iymax, int izmin, int izmax, T dxsqic, T dysqic,
    T dzsqic, T spm1, const uint32_t *__restrict__ mask, SubList<T> sub,
    int nbx, int nby, int nbdense, const int *__restrict__ skip) {
  WM_SKIP();
  int b = blockIdx.x;
  if (b < nbdense) {
    int lb = xcd_remap(b, nbdense);
    int bx = lb % nbx;
    int t = lb / nbx;
    int by = t % nby;
    int bz = t / nby;
    int j = bx * 64 + (threadIdx.x & 63);
    int iy = iymin + by * 4 + (threadIdx.x >> 6);
    int iz = izmin + bz;
    if (iy > iymax || iz > izmax) return;
    int py = iy + 1, pz = iz + 1;
    int p0 = (c + py + pz + 1) & 1;
    int px = p0 + 2 * j;
    int ix = px - 1;
    if (ix < ixmin || ix > ixmax) return;
    long long row = (long long)g.W * (py + (long long)g.PY * pz);
    long long e = row + j;
    long long ce = c * g.cs + e;
    if (mask && ((mask[ce >> 5] >> (ce & 31)) & 1u)) return;
    const T *o = phi + (1 - c) * g.cs;
    T pxm = o[row + ((px - 1) >> 1)];
    T pxp = o[row + ((px + 1) >> 1)];
    T pym = o[e - g.W];
    T pyp = o[e + g.W];
    long long pl = (long long)g.W * g.PY;
    T pzm = o[e - pl];
    T pzp = o[e + pl];
    T own = phi[ce];
    phi[ce] = rho[ce] + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic +
              (pzm + pzp) * dzsqic + spm1 * own;
    return;
  }
  int i = (b - nbdense) * 256 + threadIdx.x;
  if (i >= sub.n) return;
  int ix = sub.pos[3 * i], iy = sub.pos[3 * i + 1], iz = sub.pos[3 * i + 2];
  uint32_t m = sub.modes[i];
  const T *k = sub.k + 6 * i;
  T pxm = side_value<T>(m & 3u, k[0], phi[rb_index(g, ix - 1, iy, iz)]);
  T pxp = side_value<T>((m >> 2) & 3u, k[1], phi[rb_index(g, ix + 1, iy, iz)]);
  T pym = side_value<T>((m >> 4) & 3u, k[2], phi[rb_index(g, ix, iy - 1, iz)]);
  T pyp = side_value<T>((m >> 6) & 3u, k[3], phi[rb_index(g, ix, iy + 1, iz)]);
  T pzm = side_value<T>((m >> 8) & 3u, k[4], phi[rb_index(g, ix, iy, iz - 1)]);
  T pzp = side_value<T>((m >> 10) & 3u, k[5], phi[rb_index(g, ix, iy, iz + 1)]);
  long long ce = rb_index(g, ix, iy, iz);
  T prev = phi[ce];
  phi[ce] = (rho[ce] + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic +
             (pzm + pzp) * dzsqic) / sub.denom[i] + spm1 * prev;
}

// ---------------------------------------------------------------------------
// SOR half sweep, z-marching form.  Same arithmetic and the same sparse
// tail as k_sor_halfpass; the dense part gives each thread one column
// (colour element j of row iy) and walks it through ZC planes, so the
// other-colour values below, at and above the point stay in registers and
// each plane of the other colour is fetched from HBM once.  A wavefront is
// one row segment of 64 elements: the x-neighbour on the far side comes
// from the next lane (a shuffle), the y-neighbours from the rows beside
// it, which the neighbouring waves of the block load in the same step.
// Blocks: (row segment jseg, blockDim/64 rows, z chunk), remapped per XCD.
// ---------------------------------------------------------------------------
// One subgrid record of the sweep (condbndymg's update, as the record part
// of k_sor_zmarch below) that also compares the new value with save, copies
// it there and returns |save - new|.
template <typename S, typename T>
__device__ inline double sor_record_conv(S *__restrict__ phi, const S *__restrict__ rho,
                                         const RBGeom &g, const SubList<T> &sub, int i,
                                         T dxsqic, T dysqic, T dzsqic, T spm1,
                                         S *__restrict__ save, bool wsave) {
  int ix = sub.pos[3 * i], iy = sub.pos[3 * i + 1], iz = sub.pos[3 * i + 2];
  uint32_t m = sub.modes[i];
  const T *k = sub.k + 6 * i;
  T pxm = side_value<T>(m & 3u, k[0], St<S>::ld(phi[rb_index(g, ix - 1, iy, iz)]));
  T pxp = side_value<T>((m >> 2) & 3u, k[1], St<S>::ld(phi[rb_index(g, ix + 1, iy, iz)]));
  T pym = side_value<T>((m >> 4) & 3u, k[2], St<S>::ld(phi[rb_index(g, ix, iy - 1, iz)]));
  T pyp = side_value<T>((m >> 6) & 3u, k[3], St<S>::ld(phi[rb_index(g, ix, iy + 1, iz)]));
  T pzm = side_value<T>((m >> 8) & 3u, k[4], St<S>::ld(phi[rb_index(g, ix, iy, iz - 1)]));
  T pzp = side_value<T>((m >> 10) & 3u, k[5], St<S>::ld(phi[rb_index(g, ix, iy, iz + 1)]));
  long long ce = rb_index(g, ix, iy, iz);
  T prev = St<S>::ld(phi[ce]);
  const S nv = St<S>::st((St<S>::ld(rho[ce]) + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic +
                          (pzm + pzp) * dzsqic) / sub.denom[i] + spm1 * prev);
  phi[ce] = nv;
  const double d = fabs((double)St<S>::ld(save[ce]) - (double)St<S>::ld(nv));
  if (wsave) save[ce] = nv;
  return d > 0. ? d : 0.;
}

// One subgrid record of a half sweep (condbndymg, f3d_mgrid.F:2766-2980):
// the stencil with the record's side forms, divided by its denominator.
// The neighbours (all of the other colour) are read from ob, the point's
// value before the sweep from mb; both are phi for a sweep in place.
template <typename S, typename T>
__device__ inline void sor_record(S *__restrict__ phi, const S *ob, const S *mb,
                                  const S *__restrict__ rho, const RBGeom &g,
                                  const SubList<T> &sub, int i, T dxsqic, T dysqic, T dzsqic,
                                  T spm1) {
  int ix = sub.pos[3 * i], iy = sub.pos[3 * i + 1], iz = sub.pos[3 * i + 2];
  uint32_t m = sub.modes[i];
  const T *k = sub.k + 6 * i;
  T pxm = side_value<T>(m & 3u, k[0], St<S>::ld(ob[rb_index(g, ix - 1, iy, iz)]));
  T pxp = side_value<T>((m >> 2) & 3u, k[1], St<S>::ld(ob[rb_index(g, ix + 1, iy, iz)]));
  T pym = side_value<T>((m >> 4) & 3u, k[2], St<S>::ld(ob[rb_index(g, ix, iy - 1, iz)]));
  T pyp = side_value<T>((m >> 6) & 3u, k[3], St<S>::ld(ob[rb_index(g, ix, iy + 1, iz)]));
  T pzm = side_value<T>((m >> 8) & 3u, k[4], St<S>::ld(ob[rb_index(g, ix, iy, iz - 1)]));
  T pzp = side_value<T>((m >> 10) & 3u, k[5], St<S>::ld(ob[rb_index(g, ix, iy, iz + 1)]));
  long long ce = rb_index(g, ix, iy, iz);
  T prev = St<S>::ld(mb[ce]);
  phi[ce] = St<S>::st((St<S>::ld(rho[ce]) + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic +
                       (pzm + pzp) * dzsqic) / sub.denom[i] + spm1 * prev);
}

//
// CONV (the last half sweeps of a V-cycle, FP64): each point written is also
// compared with save, its value before the V-cycle, and copied there unless
// wsave is 0; every wavefront writes the largest |save - new| of its points
// to cpart[slot], slot = block * (wavefronts per block) + wavefront, 0 when
// it has none.
//
// src (not NULL): the sweep reads the values before it from src instead of
// phi and writes phi -- the first half sweeps of a V-cycle whose starting
// phi stays in src: colour 0 reads both colours from src, colour 1 its own
// colour from src and colour 0 (just written) from phi.
template <typename S, typename T = typename St<S>::T, bool CONV = false, bool MASK = true>
__global__ __launch_bounds__(1024) void k_sor_zmarch(
    S *__restrict__ phi, const S *__restrict__ rho, RBGeom g, int c, int ixmin,
    int ixmax, int iymin, int iymax, int izmin, int izmax, T dxsqic, T dysqic,
    T dzsqic, T spm1, const uint32_t *__restrict__ mask, SubList<T> sub,
    int j0, int nbx, int nby, int zc, int nbdense, int nbsub, S *__restrict__ save,
    double *__restrict__ cpart, const int *__restrict__ skip, const S *src, int wsave) {
  WM_SKIP();
  const int slot = (int)blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
  double cmax = 0.;
  // blocks [0, nbsub): the subgrid records, scheduled first so that they
  // overlap the dense sweep; nbsub is a multiple of 8, so the dense block
  // b runs on XCD b % 8 as xcd_remap assumes
  int b = (int)blockIdx.x - nbsub;
  if (b >= 0) {
    int lb = xcd_remap(b, nbdense);
    int bx = lb % nbx;
    int t = lb / nbx;
    int by = t % nby;
    int bz = t / nby;
    int lane = threadIdx.x & 63;
    int j = j0 + bx * 64 + lane;
    int iy = iymin + by * (int)(blockDim.x >> 6) + (threadIdx.x >> 6);
    if (iy > iymax) {  // whole wave: no shuffles are skipped mid-wave
      if (CONV && lane == 0) cpart[slot] = 0.;
      return;
    }
    int za = izmin + bz * zc;
    int zb = min(za + zc - 1, izmax);
    int py = iy + 1;
    long long pl = (long long)g.W * g.PY;
    const S *o = ((src && c == 0) ? src : phi) + (1 - c) * g.cs;
    const S *mr = (src ? src : phi) + c * g.cs;  // own colour before the sweep
    S *me = phi + c * g.cs;
    const S *rh = rho + c * g.cs;
    // Loads use jc, clamped into the padded row, so they stay in bounds for
    // lanes past the end of the row; such lanes never store.  The
    // x-neighbour loads stop at the last element a point of the range uses
    // (ix = 2j <= ixmax reads j+1), so that lanes past the range fetch no
    // line beyond it.
    int jc = min(j, g.W - 1);
    const int jhi = min((ixmax >> 1) + 1, g.W - 1);
    long long rowc = (long long)g.W * py + jc;
    long long meoff = (long long)c * g.cs + rowc;  // mask index within a plane
    // The loads of one plane (padded index pz), issued a plane ahead of use
    // and unconditionally (every lane loads its x-neighbour beyond the
    // wavefront, at a clamped element, whether or not it is an edge lane),
    // so that the compiler never has to drain the prefetch to wait for a
    // load issued on only some paths.
    struct Pl {
      T ozp, oym, oyp, own, r, edge, sv;
      uint32_t mw;
    };
    auto load = [&](int pz, Pl &q) {
      long long base = pl * pz;
      q.sv = CONV ? St<S>::ld(save[c * g.cs + rowc + base]) : (T)0.;
      q.ozp = St<S>::ld(o[rowc + base + pl]);
      q.oym = St<S>::ld(o[rowc + base - g.W]);
      q.oyp = St<S>::ld(o[rowc + base + g.W]);
      q.own = St<S>::ld(mr[rowc + base]);
      q.r = St<S>::ld(rh[rowc + base]);
      int p0 = (c + py + pz + 1) & 1;
      int jn = p0 ? j + 1 : j - 1;
      q.edge = St<S>::ld(o[base + (long long)g.W * py + min(max(jn, 0), jhi)]);
      q.mw = MASK ? mask[(meoff + base) >> 5] : 0u;
    };
    T oz_m = St<S>::ld(o[rowc + pl * za]);        // padded plane za: iz = za-1
    T oz_0 = St<S>::ld(o[rowc + pl * (za + 1)]);  // iz = za
    auto step = [&](int iz, const Pl &q) {
      int pz = iz + 1;
      int p0 = (c + py + pz + 1) & 1;
      // other-colour x-neighbours: (j-1, j) when p0 == 0, (j, j+1) when 1
      T sh = p0 ? __shfl_down(oz_0, 1, 64) : __shfl_up(oz_0, 1, 64);
      if (p0 ? (lane == 63) : (lane == 0)) sh = q.edge;
      int ix = p0 + 2 * j - 1;
      long long ce = meoff + pl * pz;
      bool active = j < g.W && ix >= ixmin && ix <= ixmax && !((q.mw >> (ce & 31)) & 1u);
      if (active) {
        T pxm = p0 ? oz_0 : sh;
        T pxp = p0 ? sh : oz_0;
        const S nv = St<S>::st(q.r + (pxm + pxp) * dxsqic + (q.oym + q.oyp) * dysqic +
                               (oz_m + q.ozp) * dzsqic + spm1 * q.own);
        me[rowc + pl * pz] = nv;
        if (CONV) {
          const double d = fabs((double)q.sv - (double)St<S>::ld(nv));
          cmax = d > cmax ? d : cmax;
          if (wsave) save[c * g.cs + rowc + pl * pz] = nv;
        }
      }
      oz_m = oz_0;
      oz_0 = q.ozp;
    };
    // two sets of loads in turn (a copy of a set would wait for its
    // loads), each refilled right after its plane is consumed, so the set
    // of the next plane is in flight meanwhile; the last planes re-read
    // plane zb+1 into the idle set (unused)
    Pl qa, qb;
    load(za + 1, qa);
    load(min(za + 2, zb + 1), qb);
    for (int iz = za; iz <= zb; iz += 2) {
      step(iz, qa);
      load(min(iz + 3, zb + 1), qa);
      if (iz + 1 > zb) break;
      step(iz + 1, qb);
      load(min(iz + 4, zb + 1), qb);
    }
    if (CONV) {
      cmax = wave_max(cmax);
      if (lane == 0) cpart[slot] = cmax;
    }
    return;
  }
  int i = (int)blockIdx.x * (int)blockDim.x + threadIdx.x;
  if (CONV) {
    // all lanes reach the reduction; lanes past the list contribute 0
    if (i < sub.n) cmax = sor_record_conv<S, T>(phi, rho, g, sub, i, dxsqic, dysqic, dzsqic,
                                                spm1, save, wsave != 0);
    cmax = wave_max(cmax);
    if ((threadIdx.x & 63) == 0) cpart[slot] = cmax;
    return;
  }
  if (i >= sub.n) return;
  sor_record<S, T>(phi, (src && c == 0) ? src : phi, src ? src : phi, rho, g, sub, i, dxsqic,
                   dysqic, dzsqic, spm1);
}

// cond_potmg (f3d_mgrid.F:2426-2472): pin interior conductor points.
template <typename S>
__global__ __launch_bounds__(256) void k_pin(S *__restrict__ phi, RBGeom g, int n,
                                             const int *__restrict__ pos,
                                             const double *__restrict__ volt,
                                             int usevolt,
                                             const int *__restrict__ skip) {
  WM_SKIP();
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  long long e = rb_index(g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  typedef typename St<S>::T T;
  phi[e] = St<S>::st(usevolt ? (T)volt[i] : (T)0.);
}

// condbndymgint (f3d_mgrid.F:2982-3092): Neumann subgrid mirrors.  The
// reference walks the even then the odd list in order and later writes win,
// so this runs as one sequential thread.  It is launched only when a
// conductor has Neumann surfaces (no configuration in use has any).
template <typename T>
__device__ inline void condbndymgint_seq(T *__restrict__ phi, const RBGeom &g, int n0, int n1,
                                         const int *__restrict__ pos0,
                                         const int *__restrict__ pos1,
                                         const double *__restrict__ dels0,
                                         const double *__restrict__ dels1, int lb0, int lb1,
                                         int lb2, int lb3, int lb4, int lb5) {
  int nx = g.nx, ny = g.ny, nz = g.nz;
  for (int p = 0; p < 2; p++) {
    int n = p ? n1 : n0;
    const int *pos = p ? pos1 : pos0;
    const double *dels = p ? dels1 : dels0;
    for (int ic = 0; ic < n; ic++) {
      int ix = pos[3 * ic], iy = pos[3 * ic + 1], iz = pos[3 * ic + 2];
      const double *d = dels + 6 * ic;
      T v = phi[rb_index(g, ix, iy, iz)];
      if (!((ix - 1 == 0 && lb0 == 0) || (ix + 1 == nx && lb1 == 0) ||
            (ix == 0 && lb0 < 1) || (ix == nx && lb1 < 1))) {
        if (0. >= d[0] && d[0] > -1.) phi[rb_index(g, ix - 1, iy, iz)] = v;
        if (0. >= d[1] && d[1] > -1.) phi[rb_index(g, ix + 1, iy, iz)] = v;
      }
      if (!((iy - 1 == 0 && lb2 == 0) || (iy + 1 == ny && lb3 == 0) ||
            (iy == 0 && lb2 < 1) || (iy == ny && lb3 < 1))) {
        if (0. >= d[2] && d[2] > -1.) phi[rb_index(g, ix, iy - 1, iz)] = v;
        if (0. >= d[3] && d[3] > -1.) phi[rb_index(g, ix, iy + 1, iz)] = v;
      }
      if (!((iz - 1 == 0 && lb4 == 0) || (iz + 1 == nz && lb5 == 0) ||
            (iz == 0 && lb4 < 1) || (iz == nz && lb5 < 1))) {
        if (0. >= d[4] && d[4] > -1.) phi[rb_index(g, ix, iy, iz - 1)] = v;
        if (0. >= d[5] && d[5] > -1.) phi[rb_index(g, ix, iy, iz + 1)] = v;
      }
    }
  }
}

template <typename T>
__global__ void k_condbndymgint(T *__restrict__ phi, RBGeom g, int n0, int n1,
                                const int *__restrict__ pos0,
                                const int *__restrict__ pos1,
                                const double *__restrict__ dels0,
                                const double *__restrict__ dels1, int lb0,
                                int lb1, int lb2, int lb3, int lb4, int lb5,
                                const int *__restrict__ skip) {
  WM_SKIP();
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  condbndymgint_seq<T>(phi, g, n0, n1, pos0, pos1, dels0, dels1, lb0, lb1, lb2, lb3, lb4, lb5);
}


// residual3d's dense part in z-marching form (the structure of
// k_sor_zmarch): blocks [0, nbdense) do colour 0, [nbdense, 2 nbdense)
// colour 1; each unmasked point of the loop range gets
// rho + sum(neighbours * d2c) - phi * mgparam (f3d_mgrid.F:3542-3546) in
// the red-black array res (same geometry as phi).  Masked points are left
// to k_residual_cond, points outside the range to k_res_shell.
template <typename S, typename TR, typename T = typename St<S>::T, bool MASK = true>
__global__ __launch_bounds__(256) void k_residual_zmarch(
    const S *__restrict__ phi, const S *__restrict__ rho, TR *__restrict__ res, RBGeom g,
    int ixmin, int ixmax, int iymin, int iymax, int izmin, int izmax, T dxsqic, T dysqic,
    T dzsqic, T mgparam, const uint32_t *__restrict__ mask, int j0, int nbx, int nby, int zc,
    int nbdense, const int *__restrict__ skip) {
  WM_SKIP();
  int c = blockIdx.x >= nbdense;
  int lb = xcd_remap(blockIdx.x - c * nbdense, nbdense);
  int bx = lb % nbx;
  int t = lb / nbx;
  int by = t % nby;
  int bz = t / nby;
  int lane = threadIdx.x & 63;
  int j = j0 + bx * 64 + lane;
  int iy = iymin + by * 4 + (threadIdx.x >> 6);
  if (iy > iymax) return;  // whole wave
  int za = izmin + bz * zc;
  int zb = min(za + zc - 1, izmax);
  int py = iy + 1;
  long long pl = (long long)g.W * g.PY;
  const S *o = phi + (1 - c) * g.cs;
  const S *me = phi + c * g.cs;
  const S *rh = rho + c * g.cs;
  TR *out = res + c * g.cs;
  // Loads use jc, clamped into the padded row, so they stay in bounds for
  // lanes past the end of the row; such lanes never store.  The
  // x-neighbour loads stop at the last element a point of the range uses.
  int jc = min(j, g.W - 1);
  const int jhi = min((ixmax >> 1) + 1, g.W - 1);
  long long rowc = (long long)g.W * py + jc;
  long long meoff = (long long)c * g.cs + rowc;
  // The loads of one plane, issued a plane ahead and unconditionally (as
  // in k_sor_zmarch: every lane loads its x-neighbour beyond the wavefront
  // at a clamped element), in two alternating sets, so that the compiler
  // never waits for the plane being prefetched.
  struct Pl {
    T ozp, oym, oyp, own, r, edge;
    uint32_t mw;
  };
  auto load = [&](int pz, Pl &q) {
    long long base = pl * pz;
    q.ozp = St<S>::ld(o[rowc + base + pl]);
    q.oym = St<S>::ld(o[rowc + base - g.W]);
    q.oyp = St<S>::ld(o[rowc + base + g.W]);
    q.own = St<S>::ld(me[rowc + base]);
    q.r = St<S>::ld(rh[rowc + base]);
    int p0 = (c + py + pz + 1) & 1;
    int jn = p0 ? j + 1 : j - 1;
    q.edge = St<S>::ld(o[base + (long long)g.W * py + min(max(jn, 0), jhi)]);
    q.mw = MASK ? mask[(meoff + base) >> 5] : 0u;
  };
  T oz_m = St<S>::ld(o[rowc + pl * za]);
  T oz_0 = St<S>::ld(o[rowc + pl * (za + 1)]);
  auto step = [&](int iz, const Pl &q) {
    int pz = iz + 1;
    int p0 = (c + py + pz + 1) & 1;
    T sh = p0 ? __shfl_down(oz_0, 1, 64) : __shfl_up(oz_0, 1, 64);
    if (p0 ? (lane == 63) : (lane == 0)) sh = q.edge;
    int ix = p0 + 2 * j - 1;
    long long ce = meoff + pl * pz;
    bool active = j < g.W && ix >= ixmin && ix <= ixmax && !((q.mw >> (ce & 31)) & 1u);
    if (active) {
      T pxm = p0 ? oz_0 : sh;
      T pxp = p0 ? sh : oz_0;
      T v = q.r + (pxm + pxp) * dxsqic + (q.oym + q.oyp) * dysqic + (oz_m + q.ozp) * dzsqic -
            q.own * mgparam;
      out[rowc + pl * pz] = St<TR>::st((typename St<TR>::T)v);
    }
    oz_m = oz_0;
    oz_0 = q.ozp;
  };
  // two sets in turn, each refilled right after its plane is consumed (as
  // in k_sor_zmarch); the last planes re-read plane zb+1 (unused)
  Pl qa, qb;
  load(za + 1, qa);
  load(min(za + 2, zb + 1), qb);
  for (int iz = za; iz <= zb; iz += 2) {
    step(iz, qa);
    load(min(iz + 3, zb + 1), qa);
    if (iz + 1 > zb) break;
    step(iz + 1, qb);
    load(min(iz + 4, zb + 1), qb);
  }
}

// residual3d's dense part, both colours in one pass: each thread owns
// element j of one row in both colour arrays -- the fine points 2j-1 (A,
// the colour whose element j is its left point) and 2j (B) -- and walks
// them through zc planes with their z-neighbours in registers, so each
// array is read once (k_residual_zmarch reads each colour's array in both
// colours' passes).  The same expression per point as k_residual_zmarch;
// masked points and points outside the range are left alone.  Loads are
// issued a plane ahead in two alternating sets, unconditionally (clamped
// elements for lanes past the row).
template <typename S, typename TR, typename T = typename St<S>::T, bool MASK = true>
__global__ __launch_bounds__(256) void k_residual_pairs(
    const S *__restrict__ phi, const S *__restrict__ rho, TR *__restrict__ res, RBGeom g,
    int ixmin, int ixmax, int iymin, int iymax, int izmin, int izmax, T dxsqic, T dysqic,
    T dzsqic, T mgparam, const uint32_t *__restrict__ mask, int j0, int j1, int nbx, int nby,
    int zc, int nbdense, const int *__restrict__ skip) {
  WM_SKIP();
  const int lb = xcd_remap(blockIdx.x, nbdense);
  const int bx = lb % nbx;
  const int t = lb / nbx;
  const int by = t % nby;
  const int bz = t / nby;
  const int lane = threadIdx.x & 63;
  const int j = j0 + bx * 64 + lane;
  const int iy = iymin + by * 4 + (threadIdx.x >> 6);
  if (iy > iymax) return;  // whole wave
  const int za = izmin + bz * zc;
  const int zb = min(za + zc - 1, izmax);
  const int py = iy + 1;
  const int W = g.W, cs = (int)g.cs, pl = g.W * g.PY;
  const int jc = min(j, W - 1);
  const bool inj = j <= j1;
  const int ixA = 2 * j - 1, ixB = 2 * j;
  const bool okA = inj && ixA >= ixmin && ixA <= ixmax;
  const bool okB = inj && ixB >= ixmin && ixB <= ixmax;
  // elements of plane pz (padded index): A's array is the colour cA
  auto eA = [&](int pz) { return ((py + pz + 1) & 1) * cs + W * (py + g.PY * pz) + jc; };
  auto eB = [&](int pz) { return (((py + pz + 1) & 1) ^ 1) * cs + W * (py + g.PY * pz) + jc; };
  struct Pl {
    T ap, bp, yam, yap, ybm, ybp, xa, xb, ra, rb;
    uint32_t ma, mb;
  };
  auto load = [&](int pz, Pl &q) {
    const int a = eA(pz), b = eB(pz);
    q.ap = St<S>::ld(phi[b + pl]);
    q.bp = St<S>::ld(phi[a + pl]);
    q.yam = St<S>::ld(phi[b - W]);
    q.yap = St<S>::ld(phi[b + W]);
    q.ybm = St<S>::ld(phi[a - W]);
    q.ybp = St<S>::ld(phi[a + W]);
    q.xa = St<S>::ld(phi[b - 1]);
    q.xb = St<S>::ld(phi[a + 1]);
    q.ra = St<S>::ld(rho[a]);
    q.rb = St<S>::ld(rho[b]);
    q.ma = MASK ? mask[a >> 5] : 0u;
    q.mb = MASK ? mask[b >> 5] : 0u;
  };
  const int pza = za + 1;
  T am = St<S>::ld(phi[eB(pza) - pl]);
  T bm = St<S>::ld(phi[eA(pza) - pl]);
  T a0 = St<S>::ld(phi[eA(pza)]);
  T b0 = St<S>::ld(phi[eB(pza)]);
  auto step = [&](int iz, const Pl &q) {
    const int pz = iz + 1;
    const int a = eA(pz), b = eB(pz);
    if (okA && !((q.ma >> (a & 31)) & 1u)) {
      T v = q.ra + (q.xa + b0) * dxsqic + (q.yam + q.yap) * dysqic + (am + q.ap) * dzsqic -
            a0 * mgparam;
      res[a] = St<TR>::st((typename St<TR>::T)v);
    }
    if (okB && !((q.mb >> (b & 31)) & 1u)) {
      T v = q.rb + (a0 + q.xb) * dxsqic + (q.ybm + q.ybp) * dysqic + (bm + q.bp) * dzsqic -
            b0 * mgparam;
      res[b] = St<TR>::st((typename St<TR>::T)v);
    }
    am = a0;
    bm = b0;
    a0 = q.ap;
    b0 = q.bp;
  };
  // two sets in turn, each refilled right after its plane is consumed (as
  // in k_sor_zmarch); the last planes re-read plane zb+1 (unused)
  Pl qa, qb;
  load(za + 1, qa);
  load(min(za + 2, zb + 1), qb);
  for (int iz = za; iz <= zb; iz += 2) {
    step(iz, qa);
    load(min(iz + 3, zb + 1), qa);
    if (iz + 1 > zb) break;
    step(iz + 1, qb);
    load(min(iz + 4, zb + 1), qb);
  }
}

// The same for FP64 phi with two elements per thread (j, j+1, j odd): the
// loads are 16-byte (element 1 of every padded row starts a 128-byte
// line), so a thread issues 14 loads for four points instead of 12 for two
// -- the texture-address units, not HBM, set the pace of these z-marches.
// Each point's residual is the same expression, with the same operands in
// the same order, as in k_residual_pairs.  Rows must be padded (W a
// multiple of 16).
template <typename TR, bool MASK = true>
__global__ __launch_bounds__(256) void k_residual_pairs2(
    const double *__restrict__ phi, const double *__restrict__ rho, TR *__restrict__ res,
    RBGeom g, int ixmin, int ixmax, int iymin, int iymax, int izmin, int izmax, double dxsqic,
    double dysqic, double dzsqic, double mgparam, const uint32_t *__restrict__ mask, int jb,
    int j1, int nbx, int nby, int zc, int nbdense, const int *__restrict__ skip) {
  WM_SKIP();
  const int lb = xcd_remap(blockIdx.x, nbdense);
  const int bx = lb % nbx;
  const int t = lb / nbx;
  const int by = t % nby;
  const int bz = t / nby;
  const int lane = threadIdx.x & 63;
  const int j = jb + 2 * (bx * 64 + lane);  // odd: elements j and j+1
  const int iy = iymin + by * 4 + (threadIdx.x >> 6);
  if (iy > iymax) return;  // whole wave
  const int za = izmin + bz * zc;
  const int zb = min(za + zc - 1, izmax);
  const int py = iy + 1;
  const int W = g.W, cs = (int)g.cs, pl = g.W * g.PY;
  const int jc = min(j, W - 3);  // loads stay in the row: j+2 <= W-1
  bool ok[4];                    // A(j), A(j+1), B(j), B(j+1)
  {
    const int ix[4] = {2 * j - 1, 2 * j + 1, 2 * j, 2 * j + 2};
    const int jj[4] = {j, j + 1, j, j + 1};
#pragma unroll
    for (int q = 0; q < 4; q++) ok[q] = jj[q] <= j1 && ix[q] >= ixmin && ix[q] <= ixmax;
  }
  auto eA = [&](int pz) { return ((py + pz + 1) & 1) * cs + W * (py + g.PY * pz) + jc; };
  auto eB = [&](int pz) { return (((py + pz + 1) & 1) ^ 1) * cs + W * (py + g.PY * pz) + jc; };
  auto ld2 = [&](int e) { return *reinterpret_cast<const double2 *>(phi + e); };
  struct Pl {
    double2 ap, bp, yam, yap, ybm, ybp, ra, rb;
    double xa, xb;
    uint32_t ma0, ma1, mb0, mb1;
  };
  auto load = [&](int pz, Pl &q) {
    const int a = eA(pz), b = eB(pz);
    q.ap = ld2(b + pl);
    q.bp = ld2(a + pl);
    q.yam = ld2(b - W);
    q.yap = ld2(b + W);
    q.ybm = ld2(a - W);
    q.ybp = ld2(a + W);
    q.ra = *reinterpret_cast<const double2 *>(rho + a);
    q.rb = *reinterpret_cast<const double2 *>(rho + b);
    q.xa = phi[b - 1];
    q.xb = phi[a + 2];
    q.ma0 = MASK ? mask[a >> 5] : 0u;
    q.ma1 = MASK ? mask[(a + 1) >> 5] : 0u;
    q.mb0 = MASK ? mask[b >> 5] : 0u;
    q.mb1 = MASK ? mask[(b + 1) >> 5] : 0u;
  };
  const int pza = za + 1;
  double2 am = ld2(eB(pza) - pl);
  double2 bm = ld2(eA(pza) - pl);
  double2 a0 = ld2(eA(pza));
  double2 b0 = ld2(eB(pza));
  auto bit = [](uint32_t w, int e) { return (w >> (e & 31)) & 1u; };
  auto st = [&](int e, double v) { res[e] = St<TR>::st((typename St<TR>::T)v); };
  auto step = [&](int iz, const Pl &q) {
    const int pz = iz + 1;
    const int a = eA(pz), b = eB(pz);
    if (ok[0] && !bit(q.ma0, a))
      st(a, q.ra.x + (q.xa + b0.x) * dxsqic + (q.yam.x + q.yap.x) * dysqic +
                (am.x + q.ap.x) * dzsqic - a0.x * mgparam);
    if (ok[1] && !bit(q.ma1, a + 1))
      st(a + 1, q.ra.y + (b0.x + b0.y) * dxsqic + (q.yam.y + q.yap.y) * dysqic +
                    (am.y + q.ap.y) * dzsqic - a0.y * mgparam);
    if (ok[2] && !bit(q.mb0, b))
      st(b, q.rb.x + (a0.x + a0.y) * dxsqic + (q.ybm.x + q.ybp.x) * dysqic +
                (bm.x + q.bp.x) * dzsqic - b0.x * mgparam);
    if (ok[3] && !bit(q.mb1, b + 1))
      st(b + 1, q.rb.y + (a0.y + q.xb) * dxsqic + (q.ybm.y + q.ybp.y) * dysqic +
                    (bm.y + q.bp.y) * dzsqic - b0.y * mgparam);
    am = a0;
    bm = b0;
    a0 = q.ap;
    b0 = q.bp;
  };
  Pl qa, qb;
  load(za + 1, qa);
  load(min(za + 2, zb + 1), qb);
  for (int iz = za; iz <= zb; iz += 2) {
    step(iz, qa);
    load(min(iz + 3, zb + 1), qa);
    if (iz + 1 > zb) break;
    step(iz + 1, qb);
    load(min(iz + 4, zb + 1), qb);
  }
}

// Zero the points of a red-black residual array outside residual3d's loop
// range (guard cells and Dirichlet faces).  Threads cover the three pairs
// of two-point-thick shells; interior points are not touched.
// Item t of the shells (t < 4 ((ny+3)(nz+3) + (nx+3)(nz+3) + (nx+3)(ny+3))).
template <typename TR>
__device__ inline void res_shell_point(TR *__restrict__ res, const RBGeom &g, int ixmin,
                                       int ixmax, int iymin, int iymax, int izmin, int izmax,
                                       long long t) {
  int nx3 = g.nx + 3, ny3 = g.ny + 3, nz3 = g.nz + 3;
  long long nxs = 4ll * ny3 * nz3, nys = 4ll * nx3 * nz3, nzs = 4ll * nx3 * ny3;
  int ix, iy, iz;
  auto pick = [](int k, int n) { return k < 2 ? k - 1 : n - 2 + k; };  // -1, 0, n, n+1
  if (t < nxs) {
    int k = (int)(t % 4);
    long long u = t / 4;
    ix = pick(k, g.nx);
    iy = (int)(u % ny3) - 1;
    iz = (int)(u / ny3) - 1;
  } else if ((t -= nxs) < nys) {
    int k = (int)(t % 4);
    long long u = t / 4;
    iy = pick(k, g.ny);
    ix = (int)(u % nx3) - 1;
    iz = (int)(u / nx3) - 1;
  } else if ((t -= nys) < nzs) {
    int k = (int)(t % 4);
    long long u = t / 4;
    iz = pick(k, g.nz);
    ix = (int)(u % nx3) - 1;
    iy = (int)(u / nx3) - 1;
  } else {
    return;
  }
  if (ix < ixmin || ix > ixmax || iy < iymin || iy > iymax || iz < izmin || iz > izmax)
    res[rb_index(g, ix, iy, iz)] = St<TR>::st((typename St<TR>::T)0.);
}

template <typename TR>
__global__ __launch_bounds__(256) void k_res_shell(TR *__restrict__ res, RBGeom g, int ixmin,
                                                   int ixmax, int iymin, int iymax, int izmin,
                                                   int izmax, const int *__restrict__ skip) {
  WM_SKIP();
  res_shell_point<TR>(res, g, ixmin, ixmax, iymin, iymax, izmin, izmax,
                      (long long)blockIdx.x * 256 + threadIdx.x);
}

// cond_potmgres + condbndyres (f3d_mgrid.F:2474-2506, 3094-3365): interior
// points get 0, active subgrid points the conductor form.  With Neumann
// surfaces the reference zeroes some neighbours' residuals while walking
// the even list then the odd list; launching one parity per call (even
// first) reproduces that order.
// One conductor record's residual (condbndyres) and the neighbour zeros
// of its Neumann sides.
template <typename S, typename TR, typename T, typename RIdx>
__device__ inline void residual_record(const S *__restrict__ phi, const S *__restrict__ rho,
                                       TR *__restrict__ res, RIdx RI, const RBGeom &g,
                                       const SubList<T> &sub, int i, T dxsqics, T dysqics,
                                       T dzsqics, T mgparam) {
  int ix = sub.pos[3 * i], iy = sub.pos[3 * i + 1], iz = sub.pos[3 * i + 2];
  uint32_t m = sub.modes[i];
  const T *k = sub.k + 6 * i;
  T pxm = side_value<T>(m & 3u, k[0], St<S>::ld(phi[rb_index(g, ix - 1, iy, iz)]));
  T pxp = side_value<T>((m >> 2) & 3u, k[1], St<S>::ld(phi[rb_index(g, ix + 1, iy, iz)]));
  T pym = side_value<T>((m >> 4) & 3u, k[2], St<S>::ld(phi[rb_index(g, ix, iy - 1, iz)]));
  T pyp = side_value<T>((m >> 6) & 3u, k[3], St<S>::ld(phi[rb_index(g, ix, iy + 1, iz)]));
  T pzm = side_value<T>((m >> 8) & 3u, k[4], St<S>::ld(phi[rb_index(g, ix, iy, iz - 1)]));
  T pzp = side_value<T>((m >> 10) & 3u, k[5], St<S>::ld(phi[rb_index(g, ix, iy, iz + 1)]));
  long long ce = rb_index(g, ix, iy, iz);
  T r = sub.fac[i] * (St<S>::ld(rho[ce]) + (pxm + pxp) * dxsqics + (pym + pyp) * dysqics +
                      (pzm + pzp) * dzsqics - St<S>::ld(phi[ce]) * mgparam * sub.denom[i]);
  res[RI(ix, iy, iz)] = St<TR>::st((typename St<TR>::T)r);
  uint32_t z = sub.zero[i];
  if (z) {
    if (z & 1u) res[RI(ix - 1, iy, iz)] = St<TR>::st((typename St<TR>::T)0.);
    if (z & 2u) res[RI(ix + 1, iy, iz)] = St<TR>::st((typename St<TR>::T)0.);
    if (z & 4u) res[RI(ix, iy - 1, iz)] = St<TR>::st((typename St<TR>::T)0.);
    if (z & 8u) res[RI(ix, iy + 1, iz)] = St<TR>::st((typename St<TR>::T)0.);
    if (z & 16u) res[RI(ix, iy, iz - 1)] = St<TR>::st((typename St<TR>::T)0.);
    if (z & 32u) res[RI(ix, iy, iz + 1)] = St<TR>::st((typename St<TR>::T)0.);
  }
}

template <typename S, typename TR, typename T = typename St<S>::T, typename RIdx = NatIdx>
__global__ __launch_bounds__(256) void k_residual_cond(
    const S *__restrict__ phi, const S *__restrict__ rho, TR *__restrict__ res, RIdx RI,
    RBGeom g, int nint, const int *__restrict__ ipos, SubList<T> sub,
    T dxsqics, T dysqics, T dzsqics, T mgparam, const int *__restrict__ skip) {
  WM_SKIP();
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nint) {
    res[RI(ipos[3 * i], ipos[3 * i + 1], ipos[3 * i + 2])] = St<TR>::st((typename St<TR>::T)0.);
    return;
  }
  i -= nint;
  if (i >= sub.n) return;
  residual_record<S, TR, T, RIdx>(phi, rho, res, RI, g, sub, i, dxsqics, dysqics, dzsqics,
                                  mgparam);
}

// Both colours' records in one launch, for record lists that zero no
// neighbour's residual (then the order of the lists does not matter).
// With ord, thread i takes record ord[i] ((colour << 31) | index: both
// lists merged in natural order, so neighbouring threads read neighbouring
// points); without it, list 0 then list 1.
template <typename S, typename TR, typename T = typename St<S>::T, typename RIdx = NatIdx>
__global__ __launch_bounds__(256) void k_residual_rec2(
    const S *__restrict__ phi, const S *__restrict__ rho, TR *__restrict__ res, RIdx RI,
    RBGeom g, SubList<T> sub0, SubList<T> sub1, T dxsqics, T dysqics, T dzsqics, T mgparam,
    const int *__restrict__ skip, const uint32_t *__restrict__ ord = nullptr) {
  WM_SKIP();
  int i = blockIdx.x * 256 + threadIdx.x;
  if (ord) {
    if (i >= sub0.n + sub1.n) return;
    const uint32_t o = ord[i];
    const int k = (int)(o & 0x7fffffffu);
    if (o >> 31)
      residual_record<S, TR, T, RIdx>(phi, rho, res, RI, g, sub1, k, dxsqics, dysqics, dzsqics,
                                      mgparam);
    else
      residual_record<S, TR, T, RIdx>(phi, rho, res, RI, g, sub0, k, dxsqics, dysqics, dzsqics,
                                      mgparam);
    return;
  }
  if (i < sub0.n)
    residual_record<S, TR, T, RIdx>(phi, rho, res, RI, g, sub0, i, dxsqics, dysqics, dzsqics,
                                    mgparam);
  else if (i - sub0.n < sub1.n)
    residual_record<S, TR, T, RIdx>(phi, rho, res, RI, g, sub1, i - sub0.n, dxsqics, dysqics,
                                    dzsqics, mgparam);
}

// ---------------------------------------------------------------------------
// applyboundaryconditions3d (f3d_mgrid.F:3847-3955), one guard cell, done
// as three launches (x, then y, then z) because the y and z parts read guard
// values the earlier parts wrote.  Each thread owns one grid line normal to
// the face, so the order of the reference's statements is kept per line.
// ---------------------------------------------------------------------------
// Line t of axis `axis` (t < (na+3)(nb+3) over the two other axes).
template <typename T, typename Idx>
__device__ inline void bc_line(T *__restrict__ u, const Idx &I, int axis, int nx, int ny, int nz,
                               int bm, int bp, int lwithdirichlet, int lzerodirichlet,
                               long long t) {
  // The line is parametrised by (a, b) over the two other axes, each
  // covering -1..n+1; ranges the reference does not touch are filtered.
  int na, nb, n;
  if (axis == 0) { na = ny; nb = nz; n = nx; }
  else if (axis == 1) { na = nx; nb = nz; n = ny; }
  else { na = nx; nb = ny; n = nz; }
  if (t >= (long long)(na + 3) * (nb + 3)) return;
  int a = (int)(t % (na + 3)) - 1;
  int b = (int)(t / (na + 3)) - 1;
  auto at = [&](int i) -> T & {
    if (axis == 0) return u[I(i, a, b)];
    if (axis == 1) return u[I(a, i, b)];
    return u[I(a, b, i)];
  };
  // The extrapolation and mirror/periodic statements cover the interior
  // range of the other axes in x (0:ny,0:nz), all of x and 0:nz in y, and
  // everything in z; zeroing covers everything.
  bool core;
  if (axis == 0) core = (a >= 0 && a <= ny && b >= 0 && b <= nz);
  else if (axis == 1) core = (b >= 0 && b <= nz);
  else core = true;
  if (lwithdirichlet) {
    if (core) {
      typedef typename St<T>::T C;
      if (bm == 0) at(-1) = St<T>::st((C)2. * St<T>::ld(at(0)) - St<T>::ld(at(1)));
      if (bp == 0) at(n + 1) = St<T>::st((C)2. * St<T>::ld(at(n)) - St<T>::ld(at(n - 1)));
    }
  } else if (lzerodirichlet) {
    if (bm == 0) at(-1) = St<T>::st((typename St<T>::T)0.);
    if (bp == 0) at(n + 1) = St<T>::st((typename St<T>::T)0.);
  }
  if (!core) return;
  if (bm == 1) at(-1) = at(1);
  if (bp == 1) at(n + 1) = at(n - 1);
  if (bm == 2 && bp == 2) {
    at(-1) = at(n - 1);
    T v0 = at(0), v1 = at(1);
    at(n) = v0;
    at(n + 1) = v1;
  }
}

template <typename T, typename Idx>
__global__ void k_bc_axis(T *__restrict__ u, Idx I, int axis, int nx, int ny,
                          int nz, int bm, int bp, int lwithdirichlet,
                          int lzerodirichlet, const int *__restrict__ skip) {
  WM_SKIP();
  bc_line<T, Idx>(u, I, axis, nx, ny, nz, bm, bp, lwithdirichlet, lzerodirichlet,
                  (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// restrict3d (f3d_mgrid.F:1593-1775): one thread per coarse node.  The
// per-axis window limits and tent weights are the reference's precomputed
// tables (wxa, ixmina, ixmaxa), built on the host.
// ---------------------------------------------------------------------------
struct RestrictAxis {
  const int *lo, *hi;  // clipped window per coarse index
  const double *w;     // (nc+1, 5) weights indexed from the unclipped start
  int cmin, cmax;      // computed coarse range
};

// Coarse node t (t < (nxc+1)(nyc+1)(nzc+1)).
template <typename TR, typename TC, typename OutIdx, typename FineIdx>
__device__ inline void restrict_point(const TR *__restrict__ res, const FineIdx &F,
                                      TC *__restrict__ out, const OutIdx &O, int nxc, int nyc,
                                      int nzc, const RestrictAxis &ax, const RestrictAxis &ay,
                                      const RestrictAxis &az, double ff, int t) {
  int n = (nxc + 1) * (nyc + 1) * (nzc + 1);
  if (t >= n) return;
  int ixc = t % (nxc + 1);
  int u = t / (nxc + 1);
  int iyc = u % (nyc + 1);
  int izc = u / (nyc + 1);
  long long oi = O(ixc, iyc, izc);
  if (ixc < ax.cmin || ixc > ax.cmax || iyc < ay.cmin || iyc > ay.cmax ||
      izc < az.cmin || izc > az.cmax) {
    out[oi] = (TC)0.;
    return;
  }
  const double *wx = ax.w + 5 * ixc, *wy = ay.w + 5 * iyc, *wz = az.w + 5 * izc;
  int x0 = ax.lo[ixc], x1 = ax.hi[ixc], y0 = ay.lo[iyc], y1 = ay.hi[iyc];
  int z0 = az.lo[izc], z1 = az.hi[izc];
  double r = 0., w = 0.;
  if (x1 - x0 == 2 && y1 - y0 == 2 && z1 - z0 == 2) {
    // 2:1 in every axis: all 27 loads issued before the (same-order) sums
    double v[27];
#pragma unroll
    for (int q = 0; q < 27; q++) v[q] = (double)res[F(x0 + q % 3, y0 + (q / 3) % 3, z0 + q / 9)];
#pragma unroll
    for (int q = 0; q < 27; q++) {
      double ww = wx[q % 3] * wy[(q / 3) % 3] * wz[q / 9];
      r = r + ww * v[q];
      w = w + ww;
    }
  } else {
    for (int iz = z0; iz <= z1; iz++)
      for (int iy = y0; iy <= y1; iy++)
        for (int ix = x0; ix <= x1; ix++) {
          double ww = wx[ix - x0] * wy[iy - y0] * wz[iz - z0];
          r = r + ww * (double)res[F(ix, iy, iz)];
          w = w + ww;
        }
  }
  if (w > 0.) {
    r = r / w;
    out[oi] = (TC)(r * ff);
  } else {
    out[oi] = (TC)0.;
  }
}

template <typename TR, typename TC, typename OutIdx, typename FineIdx = NatIdx>
__global__ __launch_bounds__(256) void k_restrict(
    const TR *__restrict__ res, FineIdx F, TC *__restrict__ out, OutIdx O,
    int nxc, int nyc, int nzc, RestrictAxis ax, RestrictAxis ay, RestrictAxis az,
    double ff, const int *__restrict__ skip) {
  WM_SKIP();
  restrict_point<TR, TC, OutIdx, FineIdx>(res, F, out, O, nxc, nyc, nzc, ax, ay, az, ff,
                                          (int)(blockIdx.x * 256 + threadIdx.x));
}

// Single-precision variant: the tent weights and sums are taken in T.
template <typename T, typename OutIdx, typename FineIdx = NatIdx>
__global__ __launch_bounds__(256) void k_restrict_t(
    const T *__restrict__ res, FineIdx F, T *__restrict__ out, OutIdx O,
    int nxc, int nyc, int nzc, RestrictAxis ax, RestrictAxis ay, RestrictAxis az,
    T ff, const int *__restrict__ skip) {
  WM_SKIP();
  int n = (nxc + 1) * (nyc + 1) * (nzc + 1);
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ixc = t % (nxc + 1);
  int u = t / (nxc + 1);
  int iyc = u % (nyc + 1);
  int izc = u / (nyc + 1);
  long long oi = O(ixc, iyc, izc);
  if (ixc < ax.cmin || ixc > ax.cmax || iyc < ay.cmin || iyc > ay.cmax ||
      izc < az.cmin || izc > az.cmax) {
    out[oi] = (T)0.;
    return;
  }
  const double *wx = ax.w + 5 * ixc, *wy = ay.w + 5 * iyc, *wz = az.w + 5 * izc;
  int x0 = ax.lo[ixc], x1 = ax.hi[ixc], y0 = ay.lo[iyc], y1 = ay.hi[iyc];
  int z0 = az.lo[izc], z1 = az.hi[izc];
  T r = 0., w = 0.;
  if (x1 - x0 == 2 && y1 - y0 == 2 && z1 - z0 == 2) {
    T v[27];
#pragma unroll
    for (int q = 0; q < 27; q++) v[q] = res[F(x0 + q % 3, y0 + (q / 3) % 3, z0 + q / 9)];
#pragma unroll
    for (int q = 0; q < 27; q++) {
      T ww = (T)wx[q % 3] * (T)wy[(q / 3) % 3] * (T)wz[q / 9];
      r = r + ww * v[q];
      w = w + ww;
    }
  } else {
    for (int iz = z0; iz <= z1; iz++)
      for (int iy = y0; iy <= y1; iy++)
        for (int ix = x0; ix <= x1; ix++) {
          T ww = (T)wx[ix - x0] * (T)wy[iy - y0] * (T)wz[iz - z0];
          r = r + ww * res[F(ix, iy, iz)];
          w = w + ww;
        }
  }
  out[oi] = w > (T)0. ? (r / w) * ff : (T)0.;
}

// Arithmetic of restrict3d: double for double residual and coarse source,
// else the coarse type (as k_restrict_t).
template <typename TR, typename TC> struct RestrictArith { typedef TC type; };
template <> struct RestrictArith<double, double> { typedef double type; };

// ---------------------------------------------------------------------------
// residual3d + restrict3d in one pass, for a level whose six boundaries are
// Dirichlet and whose restriction windows are the plain 2:1 windows
// [2j-1, 2j+1] over the coarse range [1, nc-1] in every axis.  Then every
// fine point a window reads lies in residual3d's loop range [1, n-1], the
// tent weights are (1/2, 1, 1/2) in each axis and the weight sum of every
// window is exactly 8 (all partial sums are multiples of 1/8), so the fine
// residual is computed where it is used and never stored.  Conductor points
// (mask) take their residual from res, where k_residual_cond has written
// it.
//
// One block per tile of coarse points (cxt x cyt) over a chunk of zc coarse
// planes.  For each fine plane, phase 1 computes the residuals of the
// tile's fine window into LDS: a thread owns a pair of fine points
// (ix = 2j-1, 2j: element j of both colour arrays) of one fine row, walks
// it through the planes with its z-neighbours in registers, and issues
// every load one plane ahead.  Phase 2 adds the plane's 3x3 terms to each
// coarse point's running sum in restrict3d's order (planes 2jz-1, 2jz,
// 2jz+1; within a plane rows, then columns), so the sums and the result are
// those of k_residual_zmarch + k_restrict, bit for bit.  Blocks past the
// nblocks tiles zero the coarse phi (one launch fewer per level).
// ---------------------------------------------------------------------------
//
// Every plane issues the same loads unconditionally (the last plane's
// prefetch re-reads a valid plane, unmasked lanes read a fixed element of
// res), so the compiler's wait before the LDS store covers only the
// previous plane's loads and the prefetch stays in flight across it.
template <typename S, typename TR, typename TC, bool MASK>
__global__ __launch_bounds__(1024) void k_resrestrict(
    const S *__restrict__ phi, const S *__restrict__ rho, const TR *__restrict__ res,
    TC *__restrict__ out, RBGeom g, RBGeom gc, typename St<S>::T dxsqic,
    typename St<S>::T dysqic, typename St<S>::T dzsqic, typename St<S>::T mgparam,
    const uint32_t *__restrict__ mask, int cxt, int cyt, int zc, int ntx, int nty, int nblocks,
    double ff, TC *__restrict__ zfill, long long nzfill, const int *__restrict__ skip) {
  typedef typename St<S>::T T;
  typedef typename RestrictArith<TR, TC>::type RA;
  extern __shared__ __align__(16) unsigned char wm_smem[];
  RA *lres = reinterpret_cast<RA *>(wm_smem);
  WM_SKIP();
  if ((int)blockIdx.x >= nblocks) {
    // the blocks past the tiles zero zfill[0, nzfill): the coarse level's
    // phi, its initial error, which nothing in this launch reads
    const long long stride = (long long)(gridDim.x - nblocks) * blockDim.x;
    for (long long t = (long long)(blockIdx.x - nblocks) * blockDim.x + threadIdx.x; t < nzfill;
         t += stride)
      zfill[t] = (TC)0.;
    return;
  }
  const int ncx = gc.nx - 1, ncy = gc.ny - 1, ncz = gc.nz - 1;  // coarse range [1, nc-1]
  const int lb = xcd_remap(blockIdx.x, nblocks);
  const int tx = lb % ntx, tr = lb / ntx, ty = tr % nty, tz = tr / nty;
  const int cx0 = 1 + tx * cxt, cy0 = 1 + ty * cyt, cz0 = 1 + tz * zc;
  const int mx = min(cxt, ncx - cx0 + 1), my = min(cyt, ncy - cy0 + 1);
  const int cz1 = min(cz0 + zc - 1, ncz);
  const int npx = mx + 1, nfy = 2 * my + 1;  // pairs j = cx0..cx0+mx, rows 2cy0-1..2cy1+1
  const int nfx = 2 * (cxt + 1);             // LDS row stride
  const int lsz = nfx * (2 * cyt + 1);       // LDS values per buffer
  const int cs = (int)g.cs, W = g.W, pl = g.W * g.PY;
  const int izf = 2 * cz0 - 1, izl = 2 * cz1 + 1;  // fine planes of the chunk

  // phase 1: the owned pair column
  const int pc = threadIdx.x;
  const bool own = pc < npx * nfy;
  const int pcc = own ? pc : 0;
  const int pcy = pcc / npx, pcx = pcc - pcy * npx;
  const int y = 2 * cy0 - 1 + pcy;
  const int rowoff = W * (y + 1) + cx0 + pcx;
  const int lidx = pcy * nfx + 2 * pcx;
  // loads of plane iz: the centre pair of plane iz+1 into (ca, cb), the rest
  // of plane iz into the named values
  auto elA = [&](int iz) { return ((y + iz + 1) & 1) * cs + rowoff + pl * (iz + 1); };
  auto elB = [&](int iz) { return (((y + iz + 1) & 1) ^ 1) * cs + rowoff + pl * (iz + 1); };
  auto load = [&](int iz, T &ca, T &cb, T &ya_m, T &ya_p, T &yb_m, T &yb_p, T &x_a, T &x_b,
                  T &r_a, T &r_b) {
    const int eA = elA(iz), eB = elB(iz);
    ca = St<S>::ld(phi[eB + pl]);
    cb = St<S>::ld(phi[eA + pl]);
    ya_m = St<S>::ld(phi[eB - W]);
    ya_p = St<S>::ld(phi[eB + W]);
    yb_m = St<S>::ld(phi[eA - W]);
    yb_p = St<S>::ld(phi[eA + W]);
    x_a = St<S>::ld(phi[eB - 1]);
    x_b = St<S>::ld(phi[eA + 1]);
    r_a = St<S>::ld(rho[eA]);
    r_b = St<S>::ld(rho[eB]);
  };
  // conductor points: mask words two planes ahead, their residuals (from
  // res) one plane ahead
  auto mbit = [&](uint32_t w, int e) { return (w >> (e & 31)) & 1u; };
  auto loadmask = [&](int iz, uint32_t &m_a, uint32_t &m_b) {
    if (MASK) {
      m_a = mask[elA(iz) >> 5];
      m_b = mask[elB(iz) >> 5];
    } else {
      m_a = m_b = 0u;
    }
  };
  T am = St<S>::ld(phi[elB(izf) - pl]);
  T bm = St<S>::ld(phi[elA(izf) - pl]);
  T a0 = St<S>::ld(phi[elA(izf)]);
  T b0 = St<S>::ld(phi[elB(izf)]);
  // The loads of one plane p: its values (centre pair of p+1, the rest of
  // p) and the residuals of its conductor points (addressed with p's mask
  // words, which were loaded two planes ahead, and kept raw: step() only
  // reads them where the bit is set).  Two sets alternate and the mask
  // words rotate through four slots, with the loop unrolled by four, so
  // that the next plane's loads are in flight while the current plane is
  // consumed and no loaded register is copied (a copy would wait for its
  // load).
  struct Set {
    T ap, bp, yam, yap, ybm, ybp, xa, xb, ra, rb;
    RA sa, sb;
  };
  auto issue = [&](int p, Set &d, const uint32_t *mp, uint32_t *mp2) {
    load(p, d.ap, d.bp, d.yam, d.yap, d.ybm, d.ybp, d.xa, d.xb, d.ra, d.rb);
    if (MASK) {
      const int eA = elA(p), eB = elB(p);
      d.sa = (RA)St<TR>::ld(res[mbit(mp[0], eA) ? eA : 0]);
      d.sb = (RA)St<TR>::ld(res[mbit(mp[1], eB) ? eB : 0]);
    } else {
      d.sa = d.sb = (RA)0.;
    }
    loadmask(min(p + 2, izl), mp2[0], mp2[1]);
  };
  // phase 2: the owned coarse point and its running sum
  const int q = threadIdx.x;
  const bool cown = q < mx * my;
  const int qq = cown ? q : 0;
  const int cyl = qq / mx, cxl = qq - cyl * mx;
  const int lq = 2 * cyl * nfx + 2 * cxl;
  RA rs = (RA)0.;
  const RA half = (RA)0.5, one = (RA)1.;
  auto step = [&](int iz, const Set &c, const uint32_t *mw) {
    RA *buf = lres + (iz & 1) * lsz;
    if (own) {
      RA vA, vB;
      if (mbit(mw[0], elA(iz))) {
        vA = c.sa;
      } else {
        T v = c.ra + (c.xa + b0) * dxsqic + (c.yam + c.yap) * dysqic + (am + c.ap) * dzsqic -
              a0 * mgparam;
        vA = (RA)(TR)v;
      }
      if (mbit(mw[1], elB(iz))) {
        vB = c.sb;
      } else {
        T v = c.rb + (a0 + c.xb) * dxsqic + (c.ybm + c.ybp) * dysqic + (bm + c.bp) * dzsqic -
              b0 * mgparam;
        vB = (RA)(TR)v;
      }
      buf[lidx] = vA;
      buf[lidx + 1] = vB;
    }
    // march: planes iz, iz+1 become iz-1, iz
    am = a0;
    bm = b0;
    a0 = c.ap;
    b0 = c.bp;
    __syncthreads();
    if (cown) {
      // phase 2: the plane's terms; weight of window position d (0..2) is
      // 1/2, 1, 1/2 and ww = wx * wy * wz as restrict3d forms it
      RA v[9];
#pragma unroll
      for (int d = 0; d < 9; d++) v[d] = buf[lq + (d / 3) * nfx + d % 3];
      auto add = [&](RA wz) {
#pragma unroll
        for (int d = 0; d < 9; d++) {
          RA ww = (d % 3 == 1 ? one : half) * (d / 3 == 1 ? one : half) * wz;
          rs = rs + ww * v[d];
        }
      };
      if (iz & 1) {
        const int jz = (iz + 1) >> 1;  // iz = 2(jz-1)+1 = 2jz-1
        if (jz - 1 >= cz0) {          // last plane of coarse plane jz-1
          add(half);
          out[rb_el(gc, cx0 + cxl, cy0 + cyl, jz - 1)] = (TC)((rs / (RA)8.) * (RA)ff);
        }
        if (jz <= cz1) {  // first plane of coarse plane jz
          rs = (RA)0.;
          add(half);
        }
      } else {
        add(one);
      }
    }
  };
  // plane izf + k uses set k % 2 and mask slot k % 4; the last planes
  // re-read plane izl (those values are unused)
  Set A, B;
  uint32_t M0[2], M1[2], M2[2], M3[2];
  loadmask(izf, M0[0], M0[1]);
  loadmask(min(izf + 1, izl), M1[0], M1[1]);
  issue(izf, A, M0, M2);
  // whole groups of four planes (no exit inside the body, so the waits the
  // compiler places at the loop head see one kind of predecessor), then
  // the last zero to three planes
  int iz = izf;
  for (; iz + 3 <= izl; iz += 4) {
    issue(iz + 1, B, M1, M3);
    step(iz, A, M0);
    issue(iz + 2, A, M2, M0);
    step(iz + 1, B, M1);
    issue(iz + 3, B, M3, M1);
    step(iz + 2, A, M2);
    issue(min(iz + 4, izl), A, M0, M2);
    step(iz + 3, B, M3);
  }
  if (iz <= izl) {
    issue(min(iz + 1, izl), B, M1, M3);
    step(iz, A, M0);
  }
  if (iz + 1 <= izl) {
    issue(min(iz + 2, izl), A, M2, M0);
    step(iz + 1, B, M1);
  }
  if (iz + 2 <= izl) {
    issue(min(iz + 3, izl), B, M3, M1);
    step(iz + 2, A, M2);
  }
}

// ---------------------------------------------------------------------------
// expand3d (f3d_mgrid.F:1777-1947): trilinear prolongation added to the
// fine phi, one thread per fine point in the loop range.
// ---------------------------------------------------------------------------
struct ExpandAxis {
  const int *j;     // coarse index per fine index
  const double *w;  // weight per fine index
  int lo, hi;       // fine loop range
};

template <typename T, typename FineIdx, typename CoarseIdx>
__global__ __launch_bounds__(256) void k_expand(
    T *__restrict__ phi, FineIdx F, const T *__restrict__ pc, CoarseIdx Cc,
    ExpandAxis ax, ExpandAxis ay, ExpandAxis az, const int *__restrict__ skip) {
  WM_SKIP();
  int nxr = ax.hi - ax.lo + 1, nyr = ay.hi - ay.lo + 1, nzr = az.hi - az.lo + 1;
  if (nxr <= 0 || nyr <= 0 || nzr <= 0) return;
  long long n = (long long)nxr * nyr * nzr;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ix = ax.lo + (int)(t % nxr);
  long long u = t / nxr;
  int iy = ay.lo + (int)(u % nyr);
  int iz = az.lo + (int)(u / nyr);
  int jx = ax.j[ix], jy = ay.j[iy], jz = az.j[iz];
  T wx = (T)ax.w[ix], wy = (T)ay.w[iy], wz = (T)az.w[iz];
  T one = (T)1.;
  long long fi = F(ix, iy, iz);
  phi[fi] = phi[fi] +
            (one - wx) * (one - wy) * (one - wz) * pc[Cc(jx, jy, jz)] +
            wx * (one - wy) * (one - wz) * pc[Cc(jx + 1, jy, jz)] +
            (one - wx) * wy * (one - wz) * pc[Cc(jx, jy + 1, jz)] +
            wx * wy * (one - wz) * pc[Cc(jx + 1, jy + 1, jz)] +
            (one - wx) * (one - wy) * wz * pc[Cc(jx, jy, jz + 1)] +
            wx * (one - wy) * wz * pc[Cc(jx + 1, jy, jz + 1)] +
            (one - wx) * wy * wz * pc[Cc(jx, jy + 1, jz + 1)] +
            wx * wy * wz * pc[Cc(jx + 1, jy + 1, jz + 1)];
}

// A value that is the same on every lane of the wavefront, kept in a scalar
// register.
__device__ inline int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// expand3d on red-black fine and coarse arrays.  Each thread owns element
// j of one fine row iy in both colour arrays, i.e. the fine points
// ix = 2j-1 and ix = 2j, and walks them through ZC fine planes; lanes run
// over the flattened (iy, j) range so that every lane has work.  The eight
// coarse values a fine point needs (x in {jx, jx+1}, rows jy, jy+1, planes
// jz, jz+1) stay in registers while consecutive fine planes share a coarse
// plane; the fine values of the next plane are loaded a plane ahead.  Same
// sum, term for term, as k_expand.
template <typename S, typename SC, int ZC>
__global__ __launch_bounds__(256) void k_expand_rb(
    S *__restrict__ phi, RBGeom g, const SC *__restrict__ pc, RBGeom gc, ExpandAxis ax,
    ExpandAxis ay, ExpandAxis az, int jlo, int nj, int bpc, const uint32_t *__restrict__ imask,
    const int *__restrict__ skip) {
  typedef typename St<S>::T T;
  WM_SKIP();
  const int nyr = ay.hi - ay.lo + 1;
  const int bz = blockIdx.x / bpc;
  const int t = (blockIdx.x - bz * bpc) * 256 + threadIdx.x;
  if (t >= nyr * nj) return;
  const int r = t / nj;
  const int iy = ay.lo + r;
  const int j = jlo + (t - r * nj);
  const int iz0 = az.lo + bz * ZC;
  const int nzp = min(ZC, az.hi - iz0 + 1);
  // coarse index and weights from expand3d's tables (f3d_mgrid.F:1833-1844)
  const int ixA = 2 * j - 1, ixB = 2 * j;
  const bool okA = ixA >= ax.lo && ixA <= ax.hi, okB = ixB >= ax.lo && ixB <= ax.hi;
  const int qA = ixA < ax.lo ? ax.lo : (ixA > ax.hi ? ax.hi : ixA);
  const int qB = ixB < ax.lo ? ax.lo : (ixB > ax.hi ? ax.hi : ixB);
  const int xA = ax.j[qA], xB = ax.j[qB];
  const T wxA = (T)ax.w[qA], wxB = (T)ax.w[qB];
  const int jy = ay.j[iy];
  const T wy = (T)ay.w[iy];
  const T one = (T)1.;
  const int cs = (int)g.cs;
  const int plane = g.W * g.PY;
  const int rowoff = g.W * (iy + 1) + j;
  // coarse plane jzc (a..) and jzc+1 (b..); index 0-1: x = xA, xA+1 and
  // 2-3: x = xB, xB+1; first digit: row jy or jy+1
  T a00, a01, a02, a03, a10, a11, a12, a13;
  T b00, b01, b02, b03, b10, b11, b12, b13;
#define LDC(x, y, z) ((T)St<SC>::ld(pc[rb_el(gc, (x), (y), (z))]))
  int jzc = -1000;
  // the point ix = 2j-1 of plane iz has colour (iy + iz + 1) & 1
  int cA = (iy + iz0 + 1) & 1;
  T fA = St<S>::ld(phi[cA * cs + rowoff + plane * (iz0 + 1)]);
  T fB = St<S>::ld(phi[(cA ^ 1) * cs + rowoff + plane * (iz0 + 1)]);
  for (int s = 0; s < nzp; s++) {
    const int iz = iz0 + s;
    const int jz = az.j[iz];
    const T wz = (T)az.w[iz];
    if (jz != jzc) {
      if (jz == jzc + 1) {
        a00 = b00; a01 = b01; a02 = b02; a03 = b03;
        a10 = b10; a11 = b11; a12 = b12; a13 = b13;
      } else {
        a00 = LDC(xA, jy, jz); a01 = LDC(xA + 1, jy, jz);
        a02 = LDC(xB, jy, jz); a03 = LDC(xB + 1, jy, jz);
        a10 = LDC(xA, jy + 1, jz); a11 = LDC(xA + 1, jy + 1, jz);
        a12 = LDC(xB, jy + 1, jz); a13 = LDC(xB + 1, jy + 1, jz);
      }
      b00 = LDC(xA, jy, jz + 1); b01 = LDC(xA + 1, jy, jz + 1);
      b02 = LDC(xB, jy, jz + 1); b03 = LDC(xB + 1, jy, jz + 1);
      b10 = LDC(xA, jy + 1, jz + 1); b11 = LDC(xA + 1, jy + 1, jz + 1);
      b12 = LDC(xB, jy + 1, jz + 1); b13 = LDC(xB + 1, jy + 1, jz + 1);
      jzc = jz;
    }
    const int fb = rowoff + plane * (iz + 1);
    const int eA = cA * cs + fb, eB = (cA ^ 1) * cs + fb;
    const T gA = fA, gB = fB;
    if (s + 1 < nzp) {
      fA = St<S>::ld(phi[(cA ^ 1) * cs + fb + plane]);
      fB = St<S>::ld(phi[cA * cs + fb + plane]);
    }
    T w = wxA;
    const T vA = gA + (one - w) * (one - wy) * (one - wz) * a00 +
                 w * (one - wy) * (one - wz) * a01 +
                 (one - w) * wy * (one - wz) * a10 +
                 w * wy * (one - wz) * a11 +
                 (one - w) * (one - wy) * wz * b00 +
                 w * (one - wy) * wz * b01 +
                 (one - w) * wy * wz * b10 +
                 w * wy * wz * b11;
    w = wxB;
    const T vB = gB + (one - w) * (one - wy) * (one - wz) * a02 +
                 w * (one - wy) * (one - wz) * a03 +
                 (one - w) * wy * (one - wz) * a12 +
                 w * wy * (one - wz) * a13 +
                 (one - w) * (one - wy) * wz * b02 +
                 w * (one - wy) * wz * b03 +
                 (one - w) * wy * wz * b12 +
                 w * wy * wz * b13;
    // points set in imask take the value the pin without voltages
    // (cond_potmg on an error level) gives them: zero
    const bool pA = imask && ((imask[eA >> 5] >> (eA & 31)) & 1u);
    const bool pB = imask && ((imask[eB >> 5] >> (eB & 31)) & 1u);
    if (okA) phi[eA] = St<S>::st(pA ? (T)0. : vA);
    if (okB) phi[eB] = St<S>::st(pB ? (T)0. : vB);
    cA ^= 1;
  }
#undef LDC
}

// ---------------------------------------------------------------------------
// Whole-array helpers.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ a, long long n, T v,
                                              const int *__restrict__ skip) {
  WM_SKIP();
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) a[t] = v;
}

// Zeros over a whole allocation of n elements from its (256-byte aligned)
// start: 16-byte stores, then the last elements one by one.
template <typename T>
__global__ __launch_bounds__(256) void k_zero(T *__restrict__ raw, long long n,
                                              const int *__restrict__ skip) {
  WM_SKIP();
  const long long nvec = (long long)(n * sizeof(T) / 16);
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < nvec) reinterpret_cast<uint4 *>(raw)[t] = make_uint4(0u, 0u, 0u, 0u);
  const long long tail0 = nvec * 16 / (long long)sizeof(T);
  if (t < n - tail0) {
    T z;
    __builtin_memset(&z, 0, sizeof(T));
    raw[tail0 + t] = z;
  }
}

// natural -> red-black (all elements, guards included); dst2 and dst3, when
// given, receive the same values (the solve's saved phi, the second buffer
// of the fused SOR pass).
template <typename T>
__global__ __launch_bounds__(256) void k_nat_to_rb(const double *__restrict__ src,
                                                   T *__restrict__ dst, RBGeom g,
                                                   T *__restrict__ dst2 = nullptr,
                                                   T *__restrict__ dst3 = nullptr) {
  long long sx = g.nx + 3, sy = g.ny + 3;
  long long n = sx * sy * (g.nz + 3);
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ix = (int)(t % sx) - 1;
  long long u = t / sx;
  int iy = (int)(u % sy) - 1;
  int iz = (int)(u / sy) - 1;
  const long long e = rb_index(g, ix, iy, iz);
  const T v = (T)src[t];
  dst[e] = v;
  if (dst2) dst2[e] = v;
  if (dst3) dst3[e] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_rb_to_nat(const T *__restrict__ src,
                                                   double *__restrict__ dst,
                                                   RBGeom g) {
  long long sx = g.nx + 3, sy = g.ny + 3;
  long long n = sx * sy * (g.nz + 3);
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ix = (int)(t % sx) - 1;
  long long u = t / sx;
  int iy = (int)(u % sy) - 1;
  int iz = (int)(u / sy) - 1;
  dst[t] = (double)src[rb_index(g, ix, iy, iz)];
}

// rho (0:nx)^3 natural -> red-black, scaled by reps0c (multigrid3dsolve
// f3d_mgrid.F:221-224); the caller's rho is left as the reference leaves
// it after its closing rescale (478-492): (rho*reps0c)*(1/reps0c).
__global__ __launch_bounds__(256) void k_rho_in(double *__restrict__ rho,
                                                double *__restrict__ dst, RBGeom g,
                                                double reps0c, double rinv) {
  long long sx = g.nx + 1, sy = g.ny + 1;
  long long n = sx * sy * (g.nz + 1);
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ix = (int)(t % sx);
  long long u = t / sx;
  int iy = (int)(u % sy);
  int iz = (int)(u / sy);
  double s = rho[t] * reps0c;
  dst[rb_index(g, ix, iy, iz)] = s;
  rho[t] = s * rinv;
}

// rho (0:nx)^3 natural -> red-black, unscaled (per-kernel entry points).
template <typename T>
__global__ __launch_bounds__(256) void k_rho_to_rb(const double *__restrict__ rho,
                                                   T *__restrict__ dst, RBGeom g) {
  long long sx = g.nx + 1, sy = g.ny + 1;
  long long n = sx * sy * (g.nz + 1);
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ix = (int)(t % sx);
  long long u = t / sx;
  int iy = (int)(u % sy);
  int iz = (int)(u / sy);
  dst[rb_index(g, ix, iy, iz)] = (T)rho[t];
}

// State of the V-cycle loop on the device.
struct SolveState {
  int iters;   // mgiters
  int done;    // the loop has ended (also the skip flag)
  double err;  // mgerror of the last V-cycle
};

// Most blocks of a convergence-test launch (size of the partials buffer).
#define WM_CONV_BLOCKS 2048

// The convergence test of multigrid3dsolve (f3d_mgrid.F:383-392, 313):
// mgerror = max over nodes 0..n of |phisave-phi|.  Two stages: each block
// of the test writes its max to partials[blockIdx.x] (plain stores, no
// atomics or fences), and k_converge_close, the next launch, takes the
// max of the partials, counts the V-cycle and sets `done` when the loop of
// multigrid3dsolve would stop (mgerror <= mgtol or mgiters == mgmaxiters,
// f3d_mgrid.F:383-434).
__device__ inline void block_max_to(double m, double *part, double *partials) {
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = part[0];
    for (int k = 1; k < 4; k++) b = part[k] > b ? part[k] : b;
    partials[blockIdx.x] = b;
  }
}

// One block of 1024 threads; each thread keeps four loads in flight.
__global__ __launch_bounds__(1024) void k_converge_close(const double *__restrict__ partials,
                                                         int n, SolveState *st, double mgtol,
                                                         int mgmaxiters) {
  if (st->done) return;
  __shared__ double part[16];
  double m = 0.;
  int i = threadIdx.x;
  for (; i + 3 * 1024 < n; i += 4 * 1024) {
    const double a = partials[i], b = partials[i + 1024], c = partials[i + 2048],
                 d = partials[i + 3072];
    m = a > m ? a : m;
    m = b > m ? b : m;
    m = c > m ? c : m;
    m = d > m ? d : m;
  }
  for (; i < n; i += 1024) m = partials[i] > m ? partials[i] : m;
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double err = part[0];
    for (int k = 1; k < 16; k++) err = part[k] > err ? part[k] : err;
    st->err = err;
    st->iters = st->iters + 1;
    if (!(err > mgtol && st->iters < mgmaxiters)) st->done = 1;
  }
}

// Row geometry shared by the row-walking reductions: padded row `row` of
// colour c starts at element base; its nodes 0..n are the elements whose
// ix = p0 + 2j - 1 lies in [0, nx] when the row itself is in range.
struct RowInfo {
  int base, p0;
  bool in;
};

__device__ inline RowInfo row_info(const RBGeom &g, int row) {
  int plane = g.PY * g.PZ;
  int c = row / plane;
  int rr = row - c * plane;
  int pz = rr / g.PY, py = rr - pz * g.PY;
  RowInfo r;
  r.p0 = (c + py + pz + 1) & 1;
  r.in = py >= 1 && py <= g.ny + 1 && pz >= 1 && pz <= g.nz + 1;
  r.base = c * (int)g.cs + g.W * (py + g.PY * pz);
  return r;
}

// Convergence test.  One wavefront per padded row of one colour, two rows
// in flight per wavefront, KW 64-element groups per row loaded before use.
template <typename T, int KW>
__global__ __launch_bounds__(256) void k_converge(const T *__restrict__ phi,
                                                  T *__restrict__ save, RBGeom g,
                                                  const SolveState *st,
                                                  double *__restrict__ partials, int docopy) {
  if (st->done) return;
  __shared__ double part[4];
  const int lane = threadIdx.x & 63;
  const int nrows = 2 * g.PY * g.PZ;
  const int stride = gridDim.x * 4;
  const int Wu = (g.nx + 4) >> 1;  // used elements per row (rows may be padded)
  double m = 0.;
  for (int row = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6)); row < nrows;
       row += 2 * stride) {
    RowInfo ri[2];
    bool live[2];
    T cur[2][KW], sv[2][KW];
    for (int j0 = 0; j0 < Wu; j0 += 64 * KW) {
#pragma unroll
      for (int r = 0; r < 2; r++) {
        live[r] = row + r * stride < nrows;
        ri[r] = row_info(g, live[r] ? row + r * stride : row);
#pragma unroll
        for (int k = 0; k < KW; k++) {
          int j = j0 + lane + 64 * k;
          int e = ri[r].base + (j < Wu ? j : Wu - 1);
          cur[r][k] = phi[e];
          sv[r][k] = save[e];
        }
      }
#pragma unroll
      for (int r = 0; r < 2; r++) {
#pragma unroll
        for (int k = 0; k < KW; k++) {
          int j = j0 + lane + 64 * k;
          int ix = ri[r].p0 + 2 * j - 1;
          if (live[r] && ri[r].in && ix >= 0 && ix <= g.nx) {
            double d = fabs((double)sv[r][k] - (double)cur[r][k]);
            m = d > m ? d : m;
          }
          if (docopy && live[r] && j < Wu) save[ri[r].base + j] = cur[r][k];
        }
      }
    }
  }
  block_max_to(m, part, partials);
}

// Defect-correction update fused with the convergence test: phi += e over
// every element (guards included) and max |phi_new - phi_old| over the
// nodes 0..n, reduced as in k_converge.  phi is both the saved phi of the
// iteration and the result, so no copy is kept.
template <typename TL, int KW>
__global__ __launch_bounds__(256) void k_dc_update(double *__restrict__ phi,
                                                   const TL *__restrict__ e, RBGeom g,
                                                   const SolveState *st,
                                                   double *__restrict__ partials) {
  if (st->done) return;
  __shared__ double part[4];
  const int lane = threadIdx.x & 63;
  const int nrows = 2 * g.PY * g.PZ;
  const int stride = gridDim.x * 4;
  const int Wu = (g.nx + 4) >> 1;  // used elements per row (rows may be padded)
  double m = 0.;
  for (int row = wave_uniform(blockIdx.x * 4 + (threadIdx.x >> 6)); row < nrows;
       row += 2 * stride) {
    RowInfo ri[2];
    bool live[2];
    double old[2][KW], nw[2][KW];
    for (int j0 = 0; j0 < Wu; j0 += 64 * KW) {
#pragma unroll
      for (int r = 0; r < 2; r++) {
        live[r] = row + r * stride < nrows;
        ri[r] = row_info(g, live[r] ? row + r * stride : row);
#pragma unroll
        for (int k = 0; k < KW; k++) {
          int j = j0 + lane + 64 * k;
          int q = ri[r].base + (j < Wu ? j : Wu - 1);
          old[r][k] = phi[q];
          nw[r][k] = (double)St<TL>::ld(e[q]);
        }
      }
#pragma unroll
      for (int r = 0; r < 2; r++) {
#pragma unroll
        for (int k = 0; k < KW; k++) {
          int j = j0 + lane + 64 * k;
          int ix = ri[r].p0 + 2 * j - 1;
          double v = nw[r][k] + old[r][k];
          if (live[r] && ri[r].in && ix >= 0 && ix <= g.nx) {
            double d = fabs(old[r][k] - v);
            m = d > m ? d : m;
          }
          if (live[r] && j < Wu) phi[ri[r].base + j] = v;
        }
      }
    }
  }
  block_max_to(m, part, partials);
}

// Defect-correction update: phi64 = e + phisave over the whole array
// (mgform=2 in multigrid3dsolve, f3d_mgrid.F:364-368) with e the
// low-precision correction.
template <typename TL>
__global__ __launch_bounds__(256) void k_dc_add(double *phi, const TL *e,
                                                const double *__restrict__ save,
                                                long long n,
                                                const int *__restrict__ skip) {
  WM_SKIP();
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) phi[t] = (double)St<TL>::ld(e[t]) + save[t];
}


// Initial guess by linear extrapolation in time: phi <- 2 phi - prev and
// prev <- the old phi, in one pass.
__global__ __launch_bounds__(256) void k_extrapolate(double *__restrict__ phi,
                                                     double *__restrict__ prev, long long n) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  double a = phi[t];
  phi[t] = 2. * a - prev[t];
  prev[t] = a;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_copy(const TA *__restrict__ a,
                                              TB *__restrict__ b, long long n,
                                              const int *__restrict__ skip) {
  WM_SKIP();
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) b[t] = (TB)a[t];
}
