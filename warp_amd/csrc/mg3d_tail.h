// The coarse end of the FP64 V-cycle in one launch.  Included by mg3d.hip.
//
// On the small coarse levels (a few thousand nodes) each step of vcycle
// (f3d_mgrid.F:1101-1591) -- a half sweep, the residual, the restriction,
// the expansion -- is a launch of a few microseconds whose time is launch
// latency and the z-march's serial chain, not bytes.  k_vtail runs the
// whole V-cycle from a level T down to the coarsest and back in one
// workgroup of 1024 threads: every step becomes a phase in which the
// threads stride over the step's points, and phases are separated by
// workgroup barriers.  The levels' phi, rho and res live in LDS for the
// whole launch, in the same red-black layout and element numbering as
// their HBM arrays (these levels' rows are not padded), and so do the
// read-only tables the steps use (conductor records and masks, transfer
// windows and weights): the launch gathers level T's phi and rho, the
// levels' res and the tables from HBM into LDS at its start (one list of
// items, copied with many loads in flight per thread) and writes level T's
// phi back at its end; between the two every phase reads and writes LDS
// only.
//
// Each phase evaluates its points with the arithmetic of the kernel it
// replaces (the same device functions where they exist, the same
// expressions in the same order otherwise), and the phases follow the
// launch order of the host path (sorpass, residual_rb, restrict_to,
// expand_rb and their boundary conditions), so the results are the same
// bits.
#pragma once
#include "mg3d_kernels.h"

enum {
  TAIL_PIN_AFTER = 1,  // pin (and mirror) after every half sweep
  TAIL_MIRROR = 2,     // Neumann subgrid mirrors after each pin
  TAIL_SPLIT = 4,      // interior residual zeros in a phase before the records
  TAIL_FUSEPIN = 8,    // the expansion into this level writes the first up pass's pin
  TAIL_HASMASK = 16,   // the level has conductor points (records or interior)
  TAIL_RES1 = 32,      // the residual's records and zeros can share the dense phase
};

// A record list in LDS: word offsets of the SubList arrays.
struct TailSub {
  int n, pos, modes, k, denom, fac, zero;
};
// Restriction and expansion tables in LDS (word offsets) with their ranges.
struct TailAxisR {
  int lo, hi, w, cmin, cmax;
};
struct TailAxisE {
  int j, w, lo, hi;
};

// One level of the tail, built on the host.  Offsets count 32-bit words of
// the launch's LDS; arrays of doubles start at even offsets.
struct TailLevel {
  RBGeom g;
  double *phi;              // HBM phi (the first level's is written back)
  int lphi, lrho, lres;     // phi, rho, res
  int lmask, limask, lint;  // mask and imask words, interior positions
  int sr[6];                // loop range of sorhalfpass3d
  int rr[6];                // loop range of residual3d
  int ssh[2], rsh[2];       // log2 of the padded element and row counts of the two ranges
  double sc[4];             // dxsqic, dysqic, dzsqic, 1 - mgparam of the SOR
  double rc[3];             // coefficients of the dense residual
  double rs[3];             // coefficients of the record residuals
  TailSub sor[2];           // records of the half sweeps
  TailSub rsub[2];          // records of the residual
  int nint, flags;
  int nsub[2];              // Neumann mirror lists (condbndymgint), read from HBM
  const int *subpos[2];
  const double *subdels[2];
  TailAxisR ra[3];          // restriction to the next level
  double ff;
  TailAxisE ea[3];          // expansion from the next level
};

// One HBM array gathered into LDS: words [start, start + words) of the
// gather's index space (start a multiple of TAIL_GB) go to LDS words dst.. .
struct TailItem {
  const uint32_t *src;
  int dst, words, start, pad;
};

constexpr int TAIL_MAX_ITEMS = 256;
constexpr int TAIL_GB = 32;  // words per thread of the gather (item starts are multiples)

// The launch runs a list of steps built on the host in the order the host
// path would launch them (each step's code appears once in the kernel, so
// the kernel stays small enough for the instruction cache).
enum {
  TOP_PIN = 0,       // cond_potmg (+ condbndymgint)
  TOP_HALF = 1,      // half sweep of colour arg
  TOP_BCPHI = 2,     // boundary conditions on phi
  TOP_RESID = 3,     // residual (with its boundary conditions)
  TOP_RESTRICT = 4,  // restriction to level+1 and its phi zeroed
  TOP_EXPAND = 5,    // expansion from level+1 (with its boundary conditions)
};
struct TailOp {
  int kind, level, arg, pad;
};

constexpr int TAIL_MAX_OPS = 1024;

struct TailRun {
  int nops;
  int b[6];
  double mgparam;
  int lds_words;  // LDS the launch uses (32-bit words)
  int nitems, gather_words;
};

#define TAIL_FOR(t, n) for (int t = (int)threadIdx.x; t < (int)(n); t += (int)blockDim.x)

__device__ inline bool tail_bit(const uint32_t *m, int e) { return (m[e >> 5] >> (e & 31)) & 1u; }

__device__ inline double *tail_d(uint32_t *w, int off) { return reinterpret_cast<double *>(w + off); }

__device__ inline SubList<double> tail_sub(const TailSub &s, uint32_t *w) {
  return SubList<double>{s.n, reinterpret_cast<const int *>(w + s.pos), w + s.modes,
                         tail_d(w, s.k), tail_d(w, s.denom), tail_d(w, s.fac), w + s.zero};
}

__device__ inline RestrictAxis tail_ra(const TailAxisR &a, uint32_t *w) {
  return RestrictAxis{reinterpret_cast<const int *>(w + a.lo),
                      reinterpret_cast<const int *>(w + a.hi), tail_d(w, a.w), a.cmin, a.cmax};
}

__device__ inline ExpandAxis tail_ea(const TailAxisE &a, uint32_t *w) {
  return ExpandAxis{reinterpret_cast<const int *>(w + a.j), tail_d(w, a.w), a.lo, a.hi};
}

__device__ inline const int *tail_int(const TailLevel &L, uint32_t *w) {
  return reinterpret_cast<const int *>(w + L.lint);
}

// cond_potmg without voltages (k_pin) and condbndymgint (k_condbndymgint).
__device__ inline void tail_pin(const TailLevel &L, const TailRun &R, uint32_t *w) {
  double *phi = tail_d(w, L.lphi);
  const int *ip = tail_int(L, w);
  TAIL_FOR(i, L.nint) { phi[rb_index(L.g, ip[3 * i], ip[3 * i + 1], ip[3 * i + 2])] = 0.; }
  __syncthreads();
  if (L.flags & TAIL_MIRROR) {
    if (threadIdx.x == 0)
      condbndymgint_seq<double>(phi, L.g, L.nsub[0], L.nsub[1], L.subpos[0], L.subpos[1],
                                L.subdels[0], L.subdels[1], R.b[0], R.b[1], R.b[2], R.b[3],
                                R.b[4], R.b[5]);
    __syncthreads();
  }
}

// applyboundaryconditions3d without Dirichlet guards (k_bc_axis, x then y
// then z; axes whose two bounds are Dirichlet change nothing).
__device__ inline void tail_bc(double *u, const RBGeom &g, const TailRun &R) {
  for (int a = 0; a < 3; a++) {
    if (R.b[2 * a] == 0 && R.b[2 * a + 1] == 0) continue;
    const int na = a == 0 ? g.ny : g.nx, nb = a == 2 ? g.ny : g.nz;
    TAIL_FOR(t, (na + 3) * (nb + 3)) {
      bc_line<double, RBIdx>(u, RBIdx{g}, a, g.nx, g.ny, g.nz, R.b[2 * a], R.b[2 * a + 1], 0, 0,
                             t);
    }
    __syncthreads();
  }
}

// Point (ix, iy, iz) of colour c of a loop range `r` from thread item t:
// element j of row iy of plane iz, decoded with the padded counts 2^sh[0]
// elements and 2^sh[1] rows (decoding by shifts; items past the range's
// elements, rows or planes return false).
__device__ inline bool tail_item(const int *r, const int *sh, int c, int t, int &j, int &iy,
                                 int &iz, int &ix) {
  const int j0 = (r[0] + 1) / 2, nj = (r[1] + 1) / 2 - j0 + 1;
  const int jj = t & ((1 << sh[0]) - 1);
  const int u = t >> sh[0];
  const int yy = u & ((1 << sh[1]) - 1);
  iz = r[4] + (u >> sh[1]);
  j = j0 + jj;
  iy = r[2] + yy;
  const int p0 = (c + iy + iz + 3) & 1;  // (c + py + pz + 1) & 1
  ix = p0 + 2 * j - 1;
  return jj < nj && iy <= r[3] && iz <= r[5] && ix >= r[0] && ix <= r[1];
}

// Items of a range: padded elements x padded rows x planes (0 when empty).
__device__ inline int tail_items(const int *r, const int *sh) {
  const int nj = (r[1] + 1) / 2 - (r[0] + 1) / 2 + 1;
  const int nyr = r[3] - r[2] + 1, nzr = r[5] - r[4] + 1;
  return (nj > 0 && nyr > 0 && nzr > 0) ? (nzr << (sh[0] + sh[1])) : 0;
}

// One half sweep of colour c: the dense points as k_sor_zmarch updates them
// and the colour's subgrid records (sor_record), in one loop (they write
// different points and read only the other colour and their own values).
// Like the host's halfpass, nothing (records included) runs when the y or z
// range is empty (the step list then has no pin after it either).
__device__ inline bool tail_half(const TailLevel &L, int c, uint32_t *w) {
  const RBGeom g = L.g;
  const int W = g.W, cs = (int)g.cs, pl = g.W * g.PY;
  if (L.sr[3] < L.sr[2] || L.sr[5] < L.sr[4]) return false;
  const int nd = tail_items(L.sr, L.ssh);
  double *phi = tail_d(w, L.lphi);
  const double *rho = tail_d(w, L.lrho);
  const double *o = phi + (1 - c) * cs;
  const uint32_t *mask = (L.flags & TAIL_HASMASK) ? w + L.lmask : nullptr;
  const double dxsqic = L.sc[0], dysqic = L.sc[1], dzsqic = L.sc[2], spm1 = L.sc[3];
  const SubList<double> sub = tail_sub(L.sor[c], w);
  TAIL_FOR(t, nd + sub.n) {
    if (t >= nd) {
      sor_record<double, double>(phi, phi, phi, rho, g, sub, t - nd, dxsqic, dysqic, dzsqic,
                                 spm1);
      continue;
    }
    int j, iy, iz, ix;
    if (!tail_item(L.sr, L.ssh, c, t, j, iy, iz, ix)) continue;
    const int row = W * (iy + 1 + g.PY * (iz + 1));
    const int e = row + j, ce = c * cs + e;
    if (mask && tail_bit(mask, ce)) continue;
    const int p0 = ix + 1 - 2 * j;
    const double pxm = o[row + j - 1 + p0], pxp = o[row + j + p0];
    const double pym = o[e - W], pyp = o[e + W];
    const double pzm = o[e - pl], pzp = o[e + pl];
    phi[ce] = rho[ce] + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic + (pzm + pzp) * dzsqic +
              spm1 * phi[ce];
  }
  __syncthreads();
  return true;
}

// residual3d into res (k_res_shell, k_residual_zmarch, k_residual_cond,
// boundary conditions).  The shell, the dense points, the interior zeros
// and both record lists write different points and read only phi and rho,
// so they share one phase unless a record zeroes a neighbour's residual
// (then the lists run in order, as the host launches them) or a record
// sits on an interior point (its zero must come first).
__device__ inline void tail_residual(const TailLevel &L, const TailRun &R, uint32_t *w) {
  const RBGeom g = L.g;
  const int W = g.W, cs = (int)g.cs, pl = g.W * g.PY;
  const double *phi = tail_d(w, L.lphi), *rho = tail_d(w, L.lrho);
  double *res = tail_d(w, L.lres);
  const uint32_t *mask = (L.flags & TAIL_HASMASK) ? w + L.lmask : nullptr;
  const int nsh = 4 * ((g.ny + 3) * (g.nz + 3) + (g.nx + 3) * (g.nz + 3) + (g.nx + 3) * (g.ny + 3));
  const int nd = L.rr[1] >= L.rr[0] ? tail_items(L.rr, L.rsh) : 0;
  const double dxsqic = L.rc[0], dysqic = L.rc[1], dzsqic = L.rc[2], mgparam = R.mgparam;
  const int *ip = tail_int(L, w);
  const SubList<double> s0 = tail_sub(L.rsub[0], w), s1 = tail_sub(L.rsub[1], w);
  const bool one = L.flags & TAIL_RES1;
  const bool recs = (L.flags & TAIL_HASMASK) && one;
  const int nint = recs ? L.nint : 0, n0 = recs ? s0.n : 0, n1 = recs ? s1.n : 0;
  TAIL_FOR(t, nsh + 2 * nd + nint + n0 + n1) {
    if (t < nsh) {
      res_shell_point<double>(res, g, L.rr[0], L.rr[1], L.rr[2], L.rr[3], L.rr[4], L.rr[5], t);
      continue;
    }
    int k = t - nsh;
    if (k < 2 * nd) {
      const int c = k >= nd;
      int j, iy, iz, ix;
      if (!tail_item(L.rr, L.rsh, c, k - c * nd, j, iy, iz, ix)) continue;
      const int row = W * (iy + 1 + g.PY * (iz + 1));
      const int e = row + j, ce = c * cs + e;
      if (mask && tail_bit(mask, ce)) continue;
      const int p0 = ix + 1 - 2 * j;
      const double *o = phi + (1 - c) * cs;
      const double pxm = o[row + j - 1 + p0], pxp = o[row + j + p0];
      const double pym = o[e - W], pyp = o[e + W];
      const double pzm = o[e - pl], pzp = o[e + pl];
      res[ce] = rho[ce] + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic + (pzm + pzp) * dzsqic -
                phi[ce] * mgparam;
      continue;
    }
    k -= 2 * nd;
    if (k < nint) {
      res[rb_index(g, ip[3 * k], ip[3 * k + 1], ip[3 * k + 2])] = 0.;
      continue;
    }
    k -= nint;
    if (k < n0)
      residual_record<double, double, double, RBIdx>(phi, rho, res, RBIdx{g}, g, s0, k, L.rs[0],
                                                     L.rs[1], L.rs[2], mgparam);
    else
      residual_record<double, double, double, RBIdx>(phi, rho, res, RBIdx{g}, g, s1, k - n0,
                                                     L.rs[0], L.rs[1], L.rs[2], mgparam);
  }
  __syncthreads();
  if ((L.flags & TAIL_HASMASK) && !one) {
    const bool split = L.flags & TAIL_SPLIT;
    if (split) {
      TAIL_FOR(i, L.nint) { res[rb_index(g, ip[3 * i], ip[3 * i + 1], ip[3 * i + 2])] = 0.; }
      __syncthreads();
    }
    for (int p = 0; p < 2; p++) {
      const SubList<double> sub = p ? s1 : s0;
      const int ni = p == 0 && !split ? L.nint : 0;
      const int cnt = ni + sub.n;
      if (cnt == 0) continue;
      TAIL_FOR(i, cnt) {
        if (i < ni)
          res[rb_index(g, ip[3 * i], ip[3 * i + 1], ip[3 * i + 2])] = 0.;
        else
          residual_record<double, double, double, RBIdx>(phi, rho, res, RBIdx{g}, g, sub,
                                                         i - ni, L.rs[0], L.rs[1], L.rs[2],
                                                         mgparam);
      }
      __syncthreads();
    }
  }
  tail_bc(res, g, R);
}

// restrict3d of L's residual into the source of the next level C (k_restrict),
// and C's phi set to zero (its initial error).
__device__ inline void tail_restrict(const TailLevel &L, const TailLevel &C, uint32_t *w) {
  const RBGeom gc = C.g;
  const int n = (gc.nx + 1) * (gc.ny + 1) * (gc.nz + 1);
  const double *res = tail_d(w, L.lres);
  double *rhoc = tail_d(w, C.lrho), *phic = tail_d(w, C.lphi);
  const RestrictAxis ax = tail_ra(L.ra[0], w), ay = tail_ra(L.ra[1], w), az = tail_ra(L.ra[2], w);
  TAIL_FOR(t, n) {
    restrict_point<double, double, RBIdx, RBIdx>(res, RBIdx{L.g}, rhoc, RBIdx{gc}, gc.nx, gc.ny,
                                                 gc.nz, ax, ay, az, L.ff, t);
  }
  TAIL_FOR(t, 2 * (int)gc.cs) { phic[t] = 0.; }
  __syncthreads();
}

// expand3d of C's phi added to L's (the sum of k_expand_rb, term for term),
// with the pin of the first up pass where it is fused, then the boundary
// conditions.
__device__ inline void tail_expand(const TailLevel &L, const TailLevel &C, const TailRun &R,
                                   uint32_t *w) {
  const ExpandAxis ax = tail_ea(L.ea[0], w), ay = tail_ea(L.ea[1], w), az = tail_ea(L.ea[2], w);
  const int nxr = ax.hi - ax.lo + 1, nyr = ay.hi - ay.lo + 1, nzr = az.hi - az.lo + 1;
  const bool fusepin = L.flags & TAIL_FUSEPIN;
  double *phi = tail_d(w, L.lphi);
  if (nxr <= 0 || nyr <= 0 || nzr <= 0) {
    if (fusepin) tail_pin(L, R, w);
  } else {
    const RBGeom g = L.g, gc = C.g;
    const double *pc = tail_d(w, C.lphi);
    const uint32_t *imask = w + L.limask;
    const double one = 1.;
    TAIL_FOR(t, nxr * nyr * nzr) {
      const int ix = ax.lo + t % nxr;
      const int u = t / nxr;
      const int iy = ay.lo + u % nyr, iz = az.lo + u / nyr;
      const int jx = ax.j[ix], jy = ay.j[iy], jz = az.j[iz];
      const double wx = ax.w[ix], wy = ay.w[iy], wz = az.w[iz];
      const int e = (int)rb_index(g, ix, iy, iz);
      const double v = phi[e] + (one - wx) * (one - wy) * (one - wz) * pc[rb_el(gc, jx, jy, jz)] +
                       wx * (one - wy) * (one - wz) * pc[rb_el(gc, jx + 1, jy, jz)] +
                       (one - wx) * wy * (one - wz) * pc[rb_el(gc, jx, jy + 1, jz)] +
                       wx * wy * (one - wz) * pc[rb_el(gc, jx + 1, jy + 1, jz)] +
                       (one - wx) * (one - wy) * wz * pc[rb_el(gc, jx, jy, jz + 1)] +
                       wx * (one - wy) * wz * pc[rb_el(gc, jx + 1, jy, jz + 1)] +
                       (one - wx) * wy * wz * pc[rb_el(gc, jx, jy + 1, jz + 1)] +
                       wx * wy * wz * pc[rb_el(gc, jx + 1, jy + 1, jz + 1)];
      phi[e] = (fusepin && tail_bit(imask, e)) ? 0. : v;
    }
    __syncthreads();
  }
  tail_bc(phi, L.g, R);
}

// The V-cycle from tl[0] (whose phi the caller has zeroed) to the coarsest
// level and back, as the list of steps ops.
__global__ __launch_bounds__(1024) void k_vtail(const TailLevel *__restrict__ tl,
                                                const TailItem *__restrict__ items,
                                                const TailOp *__restrict__ ops, TailRun R,
                                                const int *__restrict__ skip) {
  extern __shared__ __align__(16) uint32_t tail_lds[];
  __shared__ TailItem it[TAIL_MAX_ITEMS];
  WM_SKIP();
  uint32_t *w = tail_lds;
  // LDS starts at zero (the first level's phi is zero, the other levels'
  // phi and rho are written before they are read, and the points of res a
  // residual does not write are zero in HBM too: no kernel writes them);
  // then the gather: items start at multiples of TAIL_GB words of the
  // gather's index space, so each thread's TAIL_GB consecutive words lie in
  // one item, found by halving, and all their loads are issued before the
  // stores
  TAIL_FOR(t, R.lds_words) { w[t] = 0u; }
  TAIL_FOR(i, R.nitems) { it[i] = items[i]; }
  __syncthreads();
  for (int base = (int)threadIdx.x * TAIL_GB; base < R.gather_words;
       base += (int)blockDim.x * TAIL_GB) {
    int lo = 0;
    for (int step = TAIL_MAX_ITEMS / 2; step > 0; step >>= 1)
      if (lo + step < R.nitems && it[lo + step].start <= base) lo += step;
    const TailItem item = it[lo];
    const int q0 = base - item.start;
    uint32_t v[TAIL_GB];
#pragma unroll
    for (int k = 0; k < TAIL_GB; k++) v[k] = item.src[q0 + k < item.words ? q0 + k : 0];
#pragma unroll
    for (int k = 0; k < TAIL_GB; k++)
      if (q0 + k < item.words) w[item.dst + q0 + k] = v[k];
  }
  __syncthreads();
  for (int k = 0; k < R.nops; k++) {
    const TailOp op = ops[k];
    const TailLevel &L = tl[op.level];
    switch (op.kind) {
      case TOP_PIN: tail_pin(L, R, w); break;
      case TOP_HALF: tail_half(L, op.arg, w); break;
      case TOP_BCPHI: tail_bc(tail_d(w, L.lphi), L.g, R); break;
      case TOP_RESID: tail_residual(L, R, w); break;
      case TOP_RESTRICT: tail_restrict(L, tl[op.level + 1], w); break;
      default: tail_expand(L, tl[op.level + 1], R, w); break;
    }
  }
  const double *phi0 = tail_d(w, tl[0].lphi);
  double *out = tl[0].phi;
  TAIL_FOR(t, 2 * (int)tl[0].g.cs) { out[t] = phi0[t]; }
}
