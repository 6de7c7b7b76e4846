// Particle kernels of the electrostatic PIC step for gfx950: CIC deposit
// (setrho3ddirect1), field gather (sete_from_phi_linear3d), grid E
// (getselfe3d) and the leapfrog kicks/drift (epush3d, xpush3d), from
// pywarp90/w3d.F.  All arithmetic is FP64 in the reference's operation
// order (the library is built with -ffp-contract=off), so integer cell
// indices and every per-particle value are bit-identical to the reference;
// only the order in which deposits are summed into rho differs.
#include "wm_internal.h"
#include "condfield.h"
#include "../../include/warpmi.h"

#include "sort.h"

#include <algorithm>

namespace {

struct Mesh {
  int nx, ny, nz;
  double xmmin, ymmin, zmmin, zgrid;
  double dxi, dyi, dzi;
};

// CIC index and weights, setrho3ddirect1 (w3d.F:7883-7904): int() of the
// scaled offset (truncation), fraction by subtraction.  Returns false for a
// particle whose cell is not inside the grid (the reference does not check).
__device__ inline bool cic(const Mesh &m, double x, double y, double z, long long &ii,
                           long long &jj, long long &kk, double &u1, double &v1,
                           double &w1) {
  double sx = (x - m.xmmin) * m.dxi;
  double sy = (y - m.ymmin) * m.dyi;
  double sz = (z - m.zgrid - m.zmmin) * m.dzi;
  if (!(sx >= 0. && sx < (double)m.nx && sy >= 0. && sy < (double)m.ny && sz >= 0. &&
        sz < (double)m.nz))
    return false;
  ii = (long long)sx;
  jj = (long long)sy;
  kk = (long long)sz;
  u1 = sx - ii;
  v1 = sy - jj;
  w1 = sz - kk;
  return true;
}

// Adds the eight corner charges c[] of each lane's particle to rho.  Lanes
// that are consecutive in the wave and share a cell form a run; each run's
// charges are summed over the wave (segmented suffix sums with the run
// number as segment id, so any particle order is handled) and the run's
// first lane does the eight atomic adds.  With particles ordered by cell
// most of the atomics disappear; in random order every lane is its own run
// and the reduction is skipped.
__device__ inline void deposit_runs(double *__restrict__ rho, const Mesh &m, bool ok,
                                    long long ii, long long jj, long long kk, double c[8]) {
  int lane = threadIdx.x & 63;
  long long key = ok ? ii + (long long)(m.nx + 1) * (jj + (long long)(m.ny + 1) * kk) : -1;
  long long kprev = __shfl_up(key, 1, 64);
  bool head = ok && (lane == 0 || kprev != key);
  unsigned long long heads = __ballot(head);
  unsigned long long oks = __ballot(ok);
  if (heads != oks) {
    int seg = __popcll(heads & ((2ull << lane) - 1ull));
    if (!ok) seg = -1 - lane;
    for (int off = 1; off < 64; off <<= 1) {
      int so = __shfl_down(seg, off, 64);
      bool same = (lane + off < 64) && so == seg;
      for (int q = 0; q < 8; q++) {
        double o = __shfl_down(c[q], off, 64);
        if (same) c[q] += o;
      }
    }
  }
  if (!head) return;
  long long sx = m.nx + 1, sxy = (long long)(m.nx + 1) * (m.ny + 1);
  double *r = rho + ii + sx * jj + sxy * kk;
  atomicAdd(r, c[0]);
  atomicAdd(r + 1, c[1]);
  atomicAdd(r + sx, c[2]);
  atomicAdd(r + sx + 1, c[3]);
  atomicAdd(r + sxy, c[4]);
  atomicAdd(r + sxy + 1, c[5]);
  atomicAdd(r + sxy + sx, c[6]);
  atomicAdd(r + sxy + sx + 1, c[7]);
}

// Corner charges of setrho3ddirect1 (w3d.F:7906-7916): u0*v0*w0*g, ...
__device__ inline void corner_charges(double u1, double v1, double w1, double g, double c[8]) {
  double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
  c[0] = u0 * v0 * w0 * g;
  c[1] = u1 * v0 * w0 * g;
  c[2] = u0 * v1 * w0 * g;
  c[3] = u1 * v1 * w0 * g;
  c[4] = u0 * v0 * w1 * g;
  c[5] = u1 * v0 * w1 * g;
  c[6] = u0 * v1 * w1 * g;
  c[7] = u1 * v1 * w1 * g;
}

// Deposit in the particles' own order.  A block takes 256*DEP_K
// consecutive particles, DEP_K per thread, assigned so that in one wave
// instruction the lanes are 4*DEP_K particles apart: with particles ordered
// by cell that is about one cell apart, so the lanes' LDS atomics hit
// different nodes instead of piling onto the few nodes of one cell.  The
// block first finds the box of mesh nodes its particles touch; when it fits
// in LDS (a few hundred nodes for ordered particles) the charges are summed
// there and the box is added to rho with one atomic per touched node,
// otherwise every particle adds its eight charges to rho directly.
#define DEP_BOX 2048
#define DEP_K 4
__global__ __launch_bounds__(256) void k_setrho(double *__restrict__ rho, long long np,
                                                const double *__restrict__ xp,
                                                const double *__restrict__ yp,
                                                const double *__restrict__ zp, Mesh m,
                                                double g,
                                                unsigned long long *__restrict__ nskip) {
  __shared__ double box[DEP_BOX];
  __shared__ int lim[4][6];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const long long base = (long long)blockIdx.x * 256 * DEP_K + (long long)(l * 4 + w) * DEP_K;
  long long ii[DEP_K], jj[DEP_K], kk[DEP_K];
  double u1[DEP_K], v1[DEP_K], w1[DEP_K];
  bool ok[DEP_K];
  const int big = 1 << 30;
  int a0 = big, a1 = -big, b0 = big, b1 = -big, c0 = big, c1 = -big;
#pragma unroll
  for (int k = 0; k < DEP_K; k++) {
    long long ip = base + k;
    ok[k] = false;
    if (ip < np) {
      ok[k] = cic(m, xp[ip], yp[ip], zp[ip], ii[k], jj[k], kk[k], u1[k], v1[k], w1[k]);
      if (!ok[k] && nskip) atomicAdd(nskip, 1ull);
    }
    if (ok[k]) {
      a0 = min(a0, (int)ii[k]);
      a1 = max(a1, (int)ii[k]);
      b0 = min(b0, (int)jj[k]);
      b1 = max(b1, (int)jj[k]);
      c0 = min(c0, (int)kk[k]);
      c1 = max(c1, (int)kk[k]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    a0 = min(a0, __shfl_xor(a0, off, 64)); a1 = max(a1, __shfl_xor(a1, off, 64));
    b0 = min(b0, __shfl_xor(b0, off, 64)); b1 = max(b1, __shfl_xor(b1, off, 64));
    c0 = min(c0, __shfl_xor(c0, off, 64)); c1 = max(c1, __shfl_xor(c1, off, 64));
  }
  if (l == 0) {
    lim[w][0] = a0; lim[w][1] = a1; lim[w][2] = b0; lim[w][3] = b1; lim[w][4] = c0; lim[w][5] = c1;
  }
  __syncthreads();
  a0 = min(min(lim[0][0], lim[1][0]), min(lim[2][0], lim[3][0]));
  a1 = max(max(lim[0][1], lim[1][1]), max(lim[2][1], lim[3][1]));
  b0 = min(min(lim[0][2], lim[1][2]), min(lim[2][2], lim[3][2]));
  b1 = max(max(lim[0][3], lim[1][3]), max(lim[2][3], lim[3][3]));
  c0 = min(min(lim[0][4], lim[1][4]), min(lim[2][4], lim[3][4]));
  c1 = max(max(lim[0][5], lim[1][5]), max(lim[2][5], lim[3][5]));
  if (a1 < a0) return;  // no particle of the block is in the grid
  const long long sx = m.nx + 1, sxy = (long long)(m.nx + 1) * (m.ny + 1);
  const long long nbx = a1 - a0 + 2, nby = b1 - b0 + 2, nbz = c1 - c0 + 2;
  if (nbx * nby * nbz > DEP_BOX) {
#pragma unroll
    for (int k = 0; k < DEP_K; k++) {
      if (!ok[k]) continue;
      double c[8];
      corner_charges(u1[k], v1[k], w1[k], g, c);
      double *r = rho + ii[k] + sx * jj[k] + sxy * kk[k];
      atomicAdd(r, c[0]);
      atomicAdd(r + 1, c[1]);
      atomicAdd(r + sx, c[2]);
      atomicAdd(r + sx + 1, c[3]);
      atomicAdd(r + sxy, c[4]);
      atomicAdd(r + sxy + 1, c[5]);
      atomicAdd(r + sxy + sx, c[6]);
      atomicAdd(r + sxy + sx + 1, c[7]);
    }
    return;
  }
  const int bx = (int)nbx, bxy = (int)(nbx * nby), nb = (int)(nbx * nby * nbz);
  for (int e = threadIdx.x; e < nb; e += 256) box[e] = 0.;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < DEP_K; k++) {
    if (!ok[k]) continue;
    double c[8];
    corner_charges(u1[k], v1[k], w1[k], g, c);
    double *bp = box + (ii[k] - a0) + bx * (jj[k] - b0) + bxy * (kk[k] - c0);
    atomicAdd(bp, c[0]);
    atomicAdd(bp + 1, c[1]);
    atomicAdd(bp + bx, c[2]);
    atomicAdd(bp + bx + 1, c[3]);
    atomicAdd(bp + bxy, c[4]);
    atomicAdd(bp + bxy + 1, c[5]);
    atomicAdd(bp + bxy + bx, c[6]);
    atomicAdd(bp + bxy + bx + 1, c[7]);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nb; e += 256) {
    double v = box[e];
    if (v != 0.) {
      int ex = e % bx, ey = (e / bx) % (int)nby, ez = e / bxy;
      atomicAdd(rho + (a0 + ex) + sx * (b0 + ey) + sxy * (c0 + ez), v);
    }
  }
}

// Cell key of each particle (x fastest); out-of-grid particles get the key
// ncell and sort last.
__global__ __launch_bounds__(256) void k_keys(long long np, const double *__restrict__ xp,
                                              const double *__restrict__ yp,
                                              const double *__restrict__ zp, Mesh m,
                                              unsigned int *__restrict__ key,
                                              unsigned int *__restrict__ idx) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  long long ii, jj, kk;
  double u1, v1, w1;
  unsigned int ncell = (unsigned int)((long long)m.nx * m.ny * m.nz);
  if (cic(m, xp[ip], yp[ip], zp[ip], ii, jj, kk, u1, v1, w1))
    key[ip] = (unsigned int)(ii + (long long)m.nx * (jj + (long long)m.ny * kk));
  else
    key[ip] = ncell;
  idx[ip] = (unsigned int)ip;
}

// Block index of particlesortxyzwithcopy (top.F:5114-5120):
// int(abs(x-xmmin)/dx) + (nx+1)*(iy + (ny+1)*iz).  The reference requires
// particles inside the grid; here each index is clamped to [0, n] so that
// any input gives a valid key.
__global__ __launch_bounds__(256) void k_sortkeys(long long np, const double *__restrict__ xp,
                                                  const double *__restrict__ yp,
                                                  const double *__restrict__ zp, double xmmin,
                                                  double ymmin, double zmmin, double dx,
                                                  double dy, double dz, int nx, int ny, int nz,
                                                  unsigned int *__restrict__ key,
                                                  unsigned int *__restrict__ idx) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  double fx = fabs(xp[ip] - xmmin) / dx, fy = fabs(yp[ip] - ymmin) / dy,
         fz = fabs(zp[ip] - zmmin) / dz;
  long long ix = fx < (double)nx ? (long long)fx : nx;
  long long iy = fy < (double)ny ? (long long)fy : ny;
  long long iz = fz < (double)nz ? (long long)fz : nz;
  key[ip] = (unsigned int)(ix + (long long)(nx + 1) * (iy + (long long)(ny + 1) * iz));
  idx[ip] = (unsigned int)ip;
}

// Deposit through a permutation that orders the particles by cell.
__global__ __launch_bounds__(256) void k_setrho_sorted(
    double *__restrict__ rho, long long np, const double *__restrict__ xp,
    const double *__restrict__ yp, const double *__restrict__ zp, Mesh m, double g,
    const unsigned int *__restrict__ idx) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  long long ii = 0, jj = 0, kk = 0;
  double u1, v1, w1, c[8];
  bool ok = false;
  if (t < np) {
    long long ip = idx[t];
    ok = cic(m, xp[ip], yp[ip], zp[ip], ii, jj, kk, u1, v1, w1);
  }
  if (ok) corner_charges(u1, v1, w1, g, c);
  deposit_runs(rho, m, ok, ii, jj, kk, c);
}

// dst[c][i] = src[c][perm[i]] for up to 16 columns.
struct Columns {
  const double *src[16];
  double *dst[16];
  int n;
};

__global__ __launch_bounds__(256) void k_permute(long long np, Columns cols,
                                                 const unsigned int *__restrict__ perm) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  long long p = perm[i];
  for (int c = 0; c < cols.n; c++) cols.dst[c][i] = cols.src[c][p];
}

// sete_from_phi_linear3d (w3d.F:3963-4107): E from centred differences of
// phi at the eight cell corners, trilinear weights, in the reference's sum
// order.  Particles outside the grid get no field.
__device__ inline void gather_e(const double *__restrict__ phi, const Mesh &m, double tdxi,
                                double tdyi, double tdzi, double x, double y, double z,
                                double &ext, double &eyt, double &ezt) {
  long long i, j, k;
  double u1, v1, w1;
  if (!cic(m, x, y, z, i, j, k, u1, v1, w1)) {
    ext = eyt = ezt = 0.;
    return;
  }
  double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
  long long sx = m.nx + 3, sxy = (long long)(m.nx + 3) * (m.ny + 3);
  // p(a,b,c) = phi(i+a, j+b, k+c), phi dimensioned (-1:nx+1)^3
  const double *p0 = phi + (i + 1) + sx * (j + 1) + sxy * (k + 1);
#define P(a, b, c) p0[(a) + sx * (b) + sxy * (c)]
  ext = tdxi * (u0 * v0 * w0 * (P(-1, 0, 0) - P(1, 0, 0)) + u1 * v0 * w0 * (P(0, 0, 0) - P(2, 0, 0)) +
                u0 * v1 * w0 * (P(-1, 1, 0) - P(1, 1, 0)) + u1 * v1 * w0 * (P(0, 1, 0) - P(2, 1, 0)) +
                u0 * v0 * w1 * (P(-1, 0, 1) - P(1, 0, 1)) + u1 * v0 * w1 * (P(0, 0, 1) - P(2, 0, 1)) +
                u0 * v1 * w1 * (P(-1, 1, 1) - P(1, 1, 1)) + u1 * v1 * w1 * (P(0, 1, 1) - P(2, 1, 1)));
  eyt = tdyi * (u0 * v0 * w0 * (P(0, -1, 0) - P(0, 1, 0)) + u1 * v0 * w0 * (P(1, -1, 0) - P(1, 1, 0)) +
                u0 * v1 * w0 * (P(0, 0, 0) - P(0, 2, 0)) + u1 * v1 * w0 * (P(1, 0, 0) - P(1, 2, 0)) +
                u0 * v0 * w1 * (P(0, -1, 1) - P(0, 1, 1)) + u1 * v0 * w1 * (P(1, -1, 1) - P(1, 1, 1)) +
                u0 * v1 * w1 * (P(0, 0, 1) - P(0, 2, 1)) + u1 * v1 * w1 * (P(1, 0, 1) - P(1, 2, 1)));
  ezt = tdzi * (u0 * v0 * w0 * (P(0, 0, -1) - P(0, 0, 1)) + u1 * v0 * w0 * (P(1, 0, -1) - P(1, 0, 1)) +
                u0 * v1 * w0 * (P(0, 1, -1) - P(0, 1, 1)) + u1 * v1 * w0 * (P(1, 1, -1) - P(1, 1, 1)) +
                u0 * v0 * w1 * (P(0, 0, 0) - P(0, 0, 2)) + u1 * v0 * w1 * (P(1, 0, 0) - P(1, 0, 2)) +
                u0 * v1 * w1 * (P(0, 1, 0) - P(0, 1, 2)) + u1 * v1 * w1 * (P(1, 1, 0) - P(1, 1, 2)));
#undef P
}

// sete_from_e_linear3d (w3d.F:4362-4460): E at (x, y, z) by trilinear
// interpolation of the grid field selfe(3, 0:nx, 0:ny, 0:nz) (component
// fastest, nxguarde = 0).  A scaled position outside [0, n] gets no field
// (the reference skips the particle).  On the upper face (i = n) the term
// of weight zero reads selfe at n instead of n+1.
__device__ inline void gather_e_grid(const double *__restrict__ selfe, const Mesh &m, double x,
                                     double y, double z, double &ext, double &eyt,
                                     double &ezt) {
  double xi = (x - m.xmmin) * m.dxi;
  double yj = (y - m.ymmin) * m.dyi;
  double zk = (z - m.zgrid - m.zmmin) * m.dzi;
  // written as a negation so that a NaN position is rejected too
  if (!(xi >= 0. && xi <= (double)m.nx && yj >= 0. && yj <= (double)m.ny && zk >= 0. &&
        zk <= (double)m.nz)) {
    ext = eyt = ezt = 0.;
    return;
  }
  long long i = (long long)xi, j = (long long)yj, k = (long long)zk;
  long long i1 = i < m.nx ? i + 1 : i, j1 = j < m.ny ? j + 1 : j, k1 = k < m.nz ? k + 1 : k;
  double u1 = xi - i, v1 = yj - j, w1 = zk - k;
  double u0 = 1. - u1, v0 = 1. - v1, w0 = 1. - w1;
  long long sx = m.nx + 1, sxy = (long long)(m.nx + 1) * (m.ny + 1);
  const double *e000 = selfe + 3 * (i + sx * j + sxy * k), *e100 = selfe + 3 * (i1 + sx * j + sxy * k);
  const double *e010 = selfe + 3 * (i + sx * j1 + sxy * k), *e110 = selfe + 3 * (i1 + sx * j1 + sxy * k);
  const double *e001 = selfe + 3 * (i + sx * j + sxy * k1), *e101 = selfe + 3 * (i1 + sx * j + sxy * k1);
  const double *e011 = selfe + 3 * (i + sx * j1 + sxy * k1), *e111 = selfe + 3 * (i1 + sx * j1 + sxy * k1);
  double out[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    out[c] = u0 * v0 * w0 * e000[c] + u1 * v0 * w0 * e100[c] + u0 * v1 * w0 * e010[c] +
             u1 * v1 * w0 * e110[c] + u0 * v0 * w1 * e001[c] + u1 * v0 * w1 * e101[c] +
             u0 * v1 * w1 * e011[c] + u1 * v1 * w1 * e111[c];
  ext = out[0];
  eyt = out[1];
  ezt = out[2];
}

__global__ __launch_bounds__(256) void k_sete_from_e(const double *__restrict__ selfe,
                                                     long long np, const double *__restrict__ xp,
                                                     const double *__restrict__ yp,
                                                     const double *__restrict__ zp, Mesh m,
                                                     double *__restrict__ ex,
                                                     double *__restrict__ ey,
                                                     double *__restrict__ ez) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  double a, b, c;
  gather_e_grid(selfe, m, xp[ip], yp[ip], zp[ip], a, b, c);
  ex[ip] = ex[ip] + a;
  ey[ip] = ey[ip] + b;
  ez[ip] = ez[ip] + c;
}

__global__ __launch_bounds__(256) void k_sete(const double *__restrict__ phi, long long np,
                                              const double *__restrict__ xp,
                                              const double *__restrict__ yp,
                                              const double *__restrict__ zp, Mesh m,
                                              double tdxi, double tdyi, double tdzi,
                                              double *__restrict__ ex, double *__restrict__ ey,
                                              double *__restrict__ ez) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  double a, b, c;
  gather_e(phi, m, tdxi, tdyi, tdzi, xp[ip], yp[ip], zp[ip], a, b, c);
  ex[ip] = ex[ip] + a;
  ey[ip] = ey[ip] + b;
  ez[ip] = ez[ip] + c;
}

__global__ __launch_bounds__(256) void k_epush(long long np, double *__restrict__ ux,
                                               double *__restrict__ uy, double *__restrict__ uz,
                                               const double *__restrict__ ex,
                                               const double *__restrict__ ey,
                                               const double *__restrict__ ez, double cst) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  ux[ip] = ux[ip] + ex[ip] * cst;
  uy[ip] = uy[ip] + ey[ip] * cst;
  uz[ip] = uz[ip] + ez[ip] * cst;
}

__global__ __launch_bounds__(256) void k_xpush(long long np, double *__restrict__ x,
                                               double *__restrict__ y, double *__restrict__ z,
                                               const double *__restrict__ ux,
                                               const double *__restrict__ uy,
                                               const double *__restrict__ uz,
                                               const double *__restrict__ gaminv, double dtp) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  double gi = gaminv[ip];
  x[ip] = x[ip] + ux[ip] * gi * dtp;
  y[ip] = y[ip] + uy[ip] * gi * dtp;
  z[ip] = z[ip] + uz[ip] * gi * dtp;
}

// padvnc3d "fullv" for B = 0 (w3d.F:1816, 1956-1986): E = 0 + gather,
// two half kicks, one drift.  Particles whose new position fails the
// absorbing tests of particleboundaries3d (x <= xmmin or x >= xmmax, the
// same in y, z-zgrid <= zmmin or z-zgrid >= zmmax) are counted into
// *nlost.
// CENTER selects padvnc3d's centring (w3d.F:1848-2330, B = 0, no lattice):
// 0 "fullv": two half kicks with E(x), then the drift; 1 "halfv": one half
// kick, then the drift; 2 "synchv": one half kick, no drift.  EGRID: E
// from the precomputed grid field (efetch = 3, `phi` is then selfe) instead
// of the centred differences of phi (efetch = 1); FIELD_COND: E of
// sete3dwithconductor (f3d.lcorrectede) from the conductor field data in
// cg, efetch 1 or 3.
enum { FIELD_PHI = 0, FIELD_GRID = 1, FIELD_COND = 2 };

template <int CENTER, int FIELD = FIELD_PHI>
__global__ __launch_bounds__(256) void k_gather_push(
    const double *__restrict__ phi, long long np, double *__restrict__ x, double *__restrict__ y,
    double *__restrict__ z, double *__restrict__ ux, double *__restrict__ uy,
    double *__restrict__ uz, const double *__restrict__ gaminv, double *__restrict__ ex,
    double *__restrict__ ey, double *__restrict__ ez, Mesh m, double tdxi, double tdyi,
    double tdzi, double ch, double dt, double xmmax, double ymmax, double zmmax,
    unsigned long long *__restrict__ nlost, CondGather cg) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  bool lost = false;
  if (ip < np) {
    double xi = x[ip], yi = y[ip], zi = z[ip];
    double a, b, c;
    if (FIELD == FIELD_GRID) {
      gather_e_grid(phi, m, xi, yi, zi, a, b, c);
    } else if (FIELD == FIELD_COND) {
      if (!gather_e_cond(cg, xi, yi, zi, a, b, c)) a = b = c = 0.;
    } else {
      gather_e(phi, m, tdxi, tdyi, tdzi, xi, yi, zi, a, b, c);
    }
    double ea = 0. + a, eb = 0. + b, ec = 0. + c;
    double vx = ux[ip], vy = uy[ip], vz = uz[ip];
    vx = vx + ea * ch;
    vy = vy + eb * ch;
    vz = vz + ec * ch;
    if (CENTER == 0) {
      vx = vx + ea * ch;
      vy = vy + eb * ch;
      vz = vz + ec * ch;
    }
    ux[ip] = vx;
    uy[ip] = vy;
    uz[ip] = vz;
    if (ex) {
      ex[ip] = ea;
      ey[ip] = eb;
      ez[ip] = ec;
    }
    if (CENTER != 2) {
      double gi = gaminv[ip];
      double xn = xi + vx * gi * dt, yn = yi + vy * gi * dt, zn = zi + vz * gi * dt;
      x[ip] = xn;
      y[ip] = yn;
      z[ip] = zn;
      // the absorbing tests of particleboundaries3d (top.F:3550-3560,
      // 3779-3795): x <= xmin or x >= xmax, ..., z-zgrid >= zmax or <= zmin
      double zz = zn - m.zgrid;
      lost = xn <= m.xmmin || xn >= xmmax || yn <= m.ymmin || yn >= ymmax || zz >= zmmax ||
             zz <= m.zmmin;
    }
  }
  if (nlost) {
    unsigned long long b = __ballot(lost);
    if (b && (threadIdx.x & 63) == (__ffsll((long long)b) - 1))
      atomicAdd(nlost, (unsigned long long)__popcll(b));
  }
}

// getgrid3d (dtop.F:2818-2920), ny > 0: trilinear interpolation of a
// node-centred grid at each position inside it, in the reference's order of
// operations; f untouched outside the grid.
__global__ __launch_bounds__(256) void k_getgrid3d(long long np, const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z,
                                                   double *__restrict__ f, int nx, int ny, int nz,
                                                   const double *__restrict__ grid, double xmin,
                                                   double ymin, double zmin, double dxi,
                                                   double dyi, double dzi, int l2symtry,
                                                   int l4symtry) {
  long long ip = (long long)blockIdx.x * 256 + threadIdx.x;
  if (ip >= np) return;
  double xx, yy;
  if (l4symtry) {
    xx = fabs(x[ip] - 0.) + 0.;
    yy = fabs(y[ip] - 0.) + 0.;
  } else if (l2symtry) {
    xx = x[ip];
    yy = fabs(y[ip] - 0.) + 0.;
  } else {
    xx = x[ip];
    yy = y[ip];
  }
  double gx = (xx - xmin) * dxi, gy = (yy - ymin) * dyi, gz = (z[ip] - zmin) * dzi;
  if (!(0. <= gx && gx < nx && 0. <= gy && gy < ny && 0. <= gz && gz < nz)) return;
  long long ix = (long long)gx, iy = (long long)gy, iz = (long long)gz;
  double wx = gx - ix, wy = gy - iy, wz = gz - iz;
  const long long sx = nx + 1, sxy = (long long)(nx + 1) * (ny + 1);
  const double *g = grid + ix + sx * iy + sxy * iz;
  f[ip] = g[0] * (1. - wx) * (1. - wy) * (1. - wz) + g[1] * wx * (1. - wy) * (1. - wz) +
          g[sx] * (1. - wx) * wy * (1. - wz) + g[sx + 1] * wx * wy * (1. - wz) +
          g[sxy] * (1. - wx) * (1. - wy) * wz + g[sxy + 1] * wx * (1. - wy) * wz +
          g[sxy + sx] * (1. - wx) * wy * wz + g[sxy + sx + 1] * wx * wy * wz;
}

// getselfe3d (w3d.F:6023-6109), nxguarde = 0.
__global__ __launch_bounds__(256) void k_getselfe(const double *__restrict__ phi, int nx, int ny,
                                                  int nz, double *__restrict__ selfe,
                                                  double tdxi, double tdyi, double tdzi,
                                                  int lzero) {
  long long n = (long long)(nx + 1) * (ny + 1) * (nz + 1);
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ix = (int)(t % (nx + 1));
  long long u = t / (nx + 1);
  int iy = (int)(u % (ny + 1));
  int iz = (int)(u / (ny + 1));
  long long sx = nx + 3, sxy = (long long)(nx + 3) * (ny + 3);
  const double *p = phi + (ix + 1) + sx * (iy + 1) + sxy * (iz + 1);
  double gx = tdxi * (p[-1] - p[1]);
  double gy = tdyi * (p[-sx] - p[sx]);
  double gz = tdzi * (p[-sxy] - p[sxy]);
  double *e = selfe + 3 * t;
  if (lzero) {
    e[0] = gx;
    e[1] = gy;
    e[2] = gz;
  } else {
    e[0] = e[0] + gx;
    e[1] = e[1] + gy;
    e[2] = e[2] + gz;
  }
}

// applyrhoboundaryconditions3d (w3d.F:2660-2750) for one axis: Neumann
// faces double, periodic faces sum then copy.  One thread per face line.
__global__ void k_rhobc(double *__restrict__ rho, int axis, int nx, int ny, int nz, int bm,
                        int bp) {
  int na = axis == 0 ? ny : nx, nb = axis == 2 ? ny : nz, n = axis == 0 ? nx : axis == 1 ? ny : nz;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)(na + 1) * (nb + 1)) return;
  int a = (int)(t % (na + 1)), b = (int)(t / (na + 1));
  long long sx = nx + 1, sxy = (long long)(nx + 1) * (ny + 1);
  auto at = [&](int i) -> double & {
    if (axis == 0) return rho[i + sx * a + sxy * b];
    if (axis == 1) return rho[a + sx * i + sxy * b];
    return rho[a + sx * b + sxy * i];
  };
  if (bm == 1) at(0) = at(0) + at(0);
  if (bp == 1) at(n) = at(n) + at(n);
  if (bm == 2 || bp == 2) {
    at(0) = at(0) + at(n);
    at(n) = at(0);
  }
}

// getzmmnt's whole-beam moment sums zmmnts0(1:33) (top.F:1059-1136), in
// two deterministic passes: per-block partial sums (grid-stride loop per
// thread, then a fixed-order reduction over the block), then one block
// that adds the partials in block order.
#define NZMMNT 33
__device__ inline void zmmnt_add(double *mm, double x, double y, double z, double ux, double uy,
                                 double uz, double g, bool usegaminv) {
  const double smallpos = 1.0e-36;
  double vzi = 1. / (uz + smallpos);
  mm[0] += 1.;
  mm[1] += x;
  mm[2] += y;
  mm[3] += z;
  if (!usegaminv) {
    mm[6] += ux; mm[7] += uy; mm[8] += uz;
    mm[16] += ux * ux; mm[17] += uy * uy; mm[18] += uz * uz;
    mm[21] += z * uz; mm[24] += x * uz; mm[25] += y * uz;
    mm[26] += ux * uz; mm[27] += uy * uz;
    mm[28] += x * ux; mm[29] += y * uy; mm[30] += x * uy; mm[31] += y * ux;
    mm[32] += ux * uy;
  } else {
    mm[6] += ux * g; mm[7] += uy * g; mm[8] += uz * g;
    mm[16] += (g * ux) * (g * ux); mm[17] += (g * uy) * (g * uy); mm[18] += (g * uz) * (g * uz);
    mm[21] += z * g * uz; mm[24] += x * uz * g; mm[25] += y * uz * g;
    mm[26] += ux * uz * (g * g); mm[27] += uy * uz * (g * g);
    mm[28] += x * ux * g; mm[29] += y * uy * g; mm[30] += x * uy * g; mm[31] += y * ux * g;
    mm[32] += ux * uy * (g * g);
  }
  mm[9] += x * y;
  mm[11] += x * x;
  mm[12] += y * y;
  mm[13] += z * z;
  mm[4] += ux * vzi;
  mm[5] += uy * vzi;
  mm[10] += ux * uy * (vzi * vzi);
  mm[14] += (ux * vzi) * (ux * vzi);
  mm[15] += (uy * vzi) * (uy * vzi);
  mm[19] += x * ux * vzi;
  mm[20] += y * uy * vzi;
  mm[22] += x * uy * vzi;
  mm[23] += y * ux * vzi;
}

__global__ __launch_bounds__(256) void k_zmmnt_partial(
    long long np, const double *__restrict__ xp, const double *__restrict__ yp,
    const double *__restrict__ zp, const double *__restrict__ uxp,
    const double *__restrict__ uyp, const double *__restrict__ uzp,
    const double *__restrict__ gaminv, int usegaminv, double *__restrict__ part) {
  __shared__ double red[4][NZMMNT];
  double mm[NZMMNT];
  for (int k = 0; k < NZMMNT; k++) mm[k] = 0.;
  for (long long ip = (long long)blockIdx.x * 256 + threadIdx.x; ip < np;
       ip += (long long)gridDim.x * 256)
    zmmnt_add(mm, xp[ip], yp[ip], zp[ip], uxp[ip], uyp[ip], uzp[ip],
              usegaminv ? gaminv[ip] : 1., usegaminv != 0);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < NZMMNT; k++) {
    double v = mm[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[w][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NZMMNT) {
    int k = threadIdx.x;
    part[(long long)blockIdx.x * NZMMNT + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

// Sum of the block partials, one wavefront per moment: lane l adds blocks
// l, l+64, ... in order, then the 64 lane sums are combined by a fixed
// butterfly, so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_zmmnt_final(const double *__restrict__ part, int nb,
                                                     double *__restrict__ mm) {
  const int lane = threadIdx.x & 63;
  for (int k = threadIdx.x >> 6; k < NZMMNT; k += 4) {
    double s = 0.;
    for (int b = lane; b < nb; b += 64) s += part[(long long)b * NZMMNT + k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) mm[k] = s;
  }
}

static int zmmnt_blocks(long long np) {
  return (int)std::max<long long>(1, std::min<long long>(1024, (np + 255) / 256));
}

static int nblk(long long n) { return (int)std::max<long long>(1, cdiv(n, 256)); }

static Mesh make_mesh(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz,
                      double xmmin, double ymmin, double zmmin, double zgrid) {
  Mesh m;
  m.nx = (int)nx;
  m.ny = (int)ny;
  m.nz = (int)nz;
  m.xmmin = xmmin;
  m.ymmin = ymmin;
  m.zmmin = zmmin;
  m.zgrid = zgrid;
  m.dxi = 1. / dx;
  m.dyi = 1. / dy;
  m.dzi = 1. / dz;
  return m;
}

static void check_grid(int64_t nx, int64_t ny, int64_t nz) {
  WM_REQUIRE(nx > 0 && ny > 0 && nz > 0, "grid must have nx, ny, nz > 0");
  WM_REQUIRE((nx + 1) * (ny + 1) * (nz + 1) < (1ll << 32), "grid too large for 32-bit cell keys");
}

// Charge per unit volume of one macroparticle: g = wght*q, then /dx, /dy,
// /dz in sequence (w3d.F:7782-7785).
static double deposit_g(double q, double wght, int64_t nx, int64_t ny, int64_t nz, double dx,
                        double dy, double dz) {
  double g = wght * q;
  if (nx > 0) g = g / dx;
  if (ny > 0) g = g / dy;
  if (nz > 0) g = g / dz;
  return g;
}

struct SortedWork {
  size_t keys, idx, keys2, idx2, temp, total;
};

static SortedWork sorted_layout(int64_t np, int64_t ncell) {
  (void)ncell;
  SortedWork w{};
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t n4 = al(sizeof(unsigned int) * (size_t)std::max<int64_t>(np, 1));
  w.keys = 0;
  w.idx = n4;
  w.keys2 = 2 * n4;
  w.idx2 = 3 * n4;
  w.temp = 4 * n4;
  w.total = 4 * n4 + al(radix_sort_worksize(np));
  return w;
}

// Bits of the largest key ncell (out-of-grid particles get it).
static int key_bits(int64_t ncell) {
  int endbit = 1;
  while ((1ll << endbit) <= ncell) endbit++;
  return endbit;
}

}  // namespace

extern "C" {

int wm_setrho3d(double *rho, int64_t np, const double *xp, const double *yp, const double *zp,
                double zgrid, double q, double wght, int64_t nx, int64_t ny, int64_t nz,
                double dx, double dy, double dz, double xmmin, double ymmin, double zmmin,
                int64_t *nskipped, void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    WM_REQUIRE(np >= 0, "wm_setrho3d: np < 0");
    if (nskipped) *nskipped = 0;
    if (np == 0) return 0;
    WM_REQUIRE(rho && xp && yp && zp, "wm_setrho3d: NULL array");
    hipStream_t s = wm_stream(stream);
    Mesh m = make_mesh(nx, ny, nz, dx, dy, dz, xmmin, ymmin, zmmin, zgrid);
    double g = deposit_g(q, wght, nx, ny, nz, dx, dy, dz);
    unsigned long long *cnt = nullptr;
    if (nskipped) {
      WM_HIP(hipMallocAsync((void **)&cnt, sizeof(unsigned long long), s));
      WM_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
    }
    hipLaunchKernelGGL(k_setrho, dim3((unsigned)cdiv(np, 256 * DEP_K)), dim3(256), 0, s, rho,
                       (long long)np, xp, yp, zp, m, g, cnt);
    WM_LAUNCHED();
    if (nskipped) {
      unsigned long long h = 0;
      WM_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
      WM_HIP(hipFreeAsync(cnt, s));
      WM_HIP(hipStreamSynchronize(s));
      *nskipped = (int64_t)h;
    }
  });
}

int64_t wm_setrho3d_sorted_worksize(int64_t np, int64_t nx, int64_t ny, int64_t nz) {
  return (int64_t)sorted_layout(np, nx * ny * nz).total;
}

int wm_setrho3d_sorted(double *rho, int64_t np, const double *xp, const double *yp,
                       const double *zp, double zgrid, double q, double wght, int64_t nx,
                       int64_t ny, int64_t nz, double dx, double dy, double dz, double xmmin,
                       double ymmin, double zmmin, void *work, void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    WM_REQUIRE(np >= 0 && np < (1ll << 31), "wm_setrho3d_sorted: np out of range");
    if (np == 0) return 0;
    WM_REQUIRE(rho && xp && yp && zp && work, "wm_setrho3d_sorted: NULL array");
    hipStream_t s = wm_stream(stream);
    Mesh m = make_mesh(nx, ny, nz, dx, dy, dz, xmmin, ymmin, zmmin, zgrid);
    double g = deposit_g(q, wght, nx, ny, nz, dx, dy, dz);
    int64_t ncell = nx * ny * nz;
    SortedWork w = sorted_layout(np, ncell);
    char *base = (char *)work;
    unsigned int *keys = (unsigned int *)(base + w.keys), *idx = (unsigned int *)(base + w.idx);
    unsigned int *keys2 = (unsigned int *)(base + w.keys2), *idx2 = (unsigned int *)(base + w.idx2);
    hipLaunchKernelGGL(k_keys, dim3(nblk(np)), dim3(256), 0, s, (long long)np, xp, yp, zp, m, keys,
                       idx);
    WM_LAUNCHED();
    unsigned int *sorted_idx = nullptr;
    radix_sort_pairs(keys, idx, keys2, idx2, np, key_bits(ncell), base + w.temp, s, &sorted_idx);
    hipLaunchKernelGGL(k_setrho_sorted, dim3(nblk(np)), dim3(256), 0, s, rho, (long long)np, xp, yp,
                       zp, m, g, (const unsigned int *)sorted_idx);
    WM_LAUNCHED();
  });
}

int wm_applyrhoboundaryconditions3d(int64_t nx, int64_t ny, int64_t nz, double *rho,
                                    const int64_t *bounds, void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    WM_REQUIRE(rho && bounds, "wm_applyrhoboundaryconditions3d: NULL argument");
    hipStream_t s = wm_stream(stream);
    for (int a = 0; a < 3; a++) {
      int bm = (int)bounds[2 * a], bp = (int)bounds[2 * a + 1];
      if (bm == 0 && bp == 0) continue;
      int na = a == 0 ? (int)ny : (int)nx, nb = a == 2 ? (int)ny : (int)nz;
      hipLaunchKernelGGL(k_rhobc, dim3(nblk((long long)(na + 1) * (nb + 1))), dim3(256), 0, s, rho,
                         a, (int)nx, (int)ny, (int)nz, bm, bp);
      WM_LAUNCHED();
    }
  });
}

int wm_sete3d(const double *phi, int64_t np, const double *xp, const double *yp, const double *zp,
              double zgrid, double xmmin, double ymmin, double zmmin, double dx, double dy,
              double dz, int64_t nx, int64_t ny, int64_t nz, double *ex, double *ey, double *ez,
              void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    if (np == 0) return 0;
    WM_REQUIRE(phi && xp && yp && zp && ex && ey && ez, "wm_sete3d: NULL array");
    Mesh m = make_mesh(nx, ny, nz, dx, dy, dz, xmmin, ymmin, zmmin, zgrid);
    double tdxi = 1. / (2. * dx), tdyi = 1. / (2. * dy), tdzi = 1. / (2. * dz);
    hipLaunchKernelGGL(k_sete, dim3(nblk(np)), dim3(256), 0, wm_stream(stream), phi, (long long)np,
                       xp, yp, zp, m, tdxi, tdyi, tdzi, ex, ey, ez);
    WM_LAUNCHED();
  });
}

int wm_epush3d(int64_t np, double *uxp, double *uyp, double *uzp, const double *ex,
               const double *ey, const double *ez, double q, double m, double ddt, void *stream) {
  WM_ENTRY({
    if (np == 0) return 0;
    WM_REQUIRE(uxp && uyp && uzp && ex && ey && ez, "wm_epush3d: NULL array");
    double cst = q * ddt / m;
    hipLaunchKernelGGL(k_epush, dim3(nblk(np)), dim3(256), 0, wm_stream(stream), (long long)np, uxp,
                       uyp, uzp, ex, ey, ez, cst);
    WM_LAUNCHED();
  });
}

int wm_xpush3d(int64_t np, double *xp, double *yp, double *zp, const double *uxp,
               const double *uyp, const double *uzp, const double *gaminv, double dtp,
               void *stream) {
  WM_ENTRY({
    if (np == 0) return 0;
    WM_REQUIRE(xp && yp && zp && uxp && uyp && uzp && gaminv, "wm_xpush3d: NULL array");
    hipLaunchKernelGGL(k_xpush, dim3(nblk(np)), dim3(256), 0, wm_stream(stream), (long long)np, xp,
                       yp, zp, uxp, uyp, uzp, gaminv, dtp);
    WM_LAUNCHED();
  });
}

int wm_gather_push3d(const double *phi, int64_t np, double *xp, double *yp, double *zp,
                     double *uxp, double *uyp, double *uzp, const double *gaminv, double *ex,
                     double *ey, double *ez, double zgrid, double xmmin, double ymmin,
                     double zmmin, double dx, double dy, double dz, int64_t nx, int64_t ny,
                     int64_t nz, double q, double m, double dt, double xmmax, double ymmax,
                     double zmmax, int64_t *nlost, void *stream) {
  return wm_padvnc3d(phi, np, xp, yp, zp, uxp, uyp, uzp, gaminv, ex, ey, ez, zgrid, xmmin, ymmin,
                     zmmin, dx, dy, dz, nx, ny, nz, q, m, dt, xmmax, ymmax, zmmax, nlost,
                     WM_CENTER_FULLV, stream);
}

}  // extern "C"

template <int FIELD>
static int padvnc3d_launch(const double *phi, int64_t np, double *xp, double *yp, double *zp,
                           double *uxp, double *uyp, double *uzp, const double *gaminv,
                           double *ex, double *ey, double *ez, double zgrid, double xmmin,
                           double ymmin, double zmmin, double dx, double dy, double dz,
                           int64_t nx, int64_t ny, int64_t nz, double q, double m, double dt,
                           double xmmax, double ymmax, double zmmax, int64_t *nlost,
                           int64_t center, const CondGather &cg, void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    WM_REQUIRE(center == WM_CENTER_FULLV || center == WM_CENTER_HALFV ||
                   center == WM_CENTER_SYNCHV,
               "wm_padvnc3d: center must be WM_CENTER_FULLV, _HALFV or _SYNCHV");
    if (np == 0) return 0;
    WM_REQUIRE((phi || FIELD == FIELD_COND) && xp && yp && zp && uxp && uyp && uzp && gaminv,
               "wm_padvnc3d: NULL array");
    WM_REQUIRE((ex && ey && ez) || (!ex && !ey && !ez), "wm_padvnc3d: give all or none of ex, ey, ez");
    Mesh me = make_mesh(nx, ny, nz, dx, dy, dz, xmmin, ymmin, zmmin, zgrid);
    double tdxi = 1. / (2. * dx), tdyi = 1. / (2. * dy), tdzi = 1. / (2. * dz);
    // fulldt_s = dt*ndts*dtscale, halfdt_s = 0.5*fulldt_s (w3d.F:1747-1748);
    // epush3d's const = q*ddt/m (w3d.F:2765).
    double fulldt = dt * 1 * 1.;
    double halfdt = 0.5 * fulldt;
    double ch = q * halfdt / m;
    unsigned long long *nl = (unsigned long long *)nlost;
    if (center == WM_CENTER_FULLV)
      hipLaunchKernelGGL((k_gather_push<0, FIELD>), dim3(nblk(np)), dim3(256), 0,
                         wm_stream(stream), phi, (long long)np, xp, yp, zp, uxp, uyp, uzp, gaminv,
                         ex, ey, ez, me, tdxi, tdyi, tdzi, ch, fulldt, xmmax, ymmax, zmmax, nl, cg);
    else if (center == WM_CENTER_HALFV)
      hipLaunchKernelGGL((k_gather_push<1, FIELD>), dim3(nblk(np)), dim3(256), 0,
                         wm_stream(stream), phi, (long long)np, xp, yp, zp, uxp, uyp, uzp, gaminv,
                         ex, ey, ez, me, tdxi, tdyi, tdzi, ch, fulldt, xmmax, ymmax, zmmax, nl, cg);
    else
      hipLaunchKernelGGL((k_gather_push<2, FIELD>), dim3(nblk(np)), dim3(256), 0,
                         wm_stream(stream), phi, (long long)np, xp, yp, zp, uxp, uyp, uzp, gaminv,
                         ex, ey, ez, me, tdxi, tdyi, tdzi, ch, fulldt, xmmax, ymmax, zmmax, nl, cg);
    WM_LAUNCHED();
  });
}

extern "C" {

int wm_padvnc3d(const double *phi, int64_t np, double *xp, double *yp, double *zp,
                double *uxp, double *uyp, double *uzp, const double *gaminv, double *ex,
                double *ey, double *ez, double zgrid, double xmmin, double ymmin, double zmmin,
                double dx, double dy, double dz, int64_t nx, int64_t ny, int64_t nz, double q,
                double m, double dt, double xmmax, double ymmax, double zmmax, int64_t *nlost,
                int64_t center, void *stream) {
  return padvnc3d_launch<FIELD_PHI>(phi, np, xp, yp, zp, uxp, uyp, uzp, gaminv, ex, ey, ez,
                                    zgrid, xmmin, ymmin, zmmin, dx, dy, dz, nx, ny, nz, q, m, dt,
                                    xmmax, ymmax, zmmax, nlost, center, CondGather{}, stream);
}

int wm_padvnc3d_selfe(const double *selfe, int64_t np, double *xp, double *yp, double *zp,
                      double *uxp, double *uyp, double *uzp, const double *gaminv, double *ex,
                      double *ey, double *ez, double zgrid, double xmmin, double ymmin,
                      double zmmin, double dx, double dy, double dz, int64_t nx, int64_t ny,
                      int64_t nz, double q, double m, double dt, double xmmax, double ymmax,
                      double zmmax, int64_t *nlost, int64_t center, void *stream) {
  return padvnc3d_launch<FIELD_GRID>(selfe, np, xp, yp, zp, uxp, uyp, uzp, gaminv, ex, ey, ez,
                                     zgrid, xmmin, ymmin, zmmin, dx, dy, dz, nx, ny, nz, q, m, dt,
                                     xmmax, ymmax, zmmax, nlost, center, CondGather{}, stream);
}

int wm_padvnc3d_conductor(const wm_condfield *cf, const double *phi, const double *selfe,
                          int64_t efetch, int64_t np, double *xp, double *yp, double *zp,
                          double *uxp, double *uyp, double *uzp, const double *gaminv, double *ex,
                          double *ey, double *ez, double zgrid, double xmmin, double ymmin,
                          double zmmin, double dx, double dy, double dz, double q, double m,
                          double dt, double xmmax, double ymmax, double zmmax, int64_t *nlost,
                          int64_t center, void *stream) {
  CondGather cg;
  int rc = [&]() -> int {
    WM_ENTRY({
      cg = wm_condfield_gather(cf, phi, selfe, efetch, zgrid, xmmin, ymmin, zmmin, dx, dy, dz);
    });
  }();
  if (rc != 0) return rc;
  return padvnc3d_launch<FIELD_COND>(phi, np, xp, yp, zp, uxp, uyp, uzp, gaminv, ex, ey, ez,
                                     zgrid, xmmin, ymmin, zmmin, dx, dy, dz, cg.v.nx, cg.v.ny,
                                     cg.v.nz, q, m, dt, xmmax, ymmax, zmmax, nlost, center, cg,
                                     stream);
}

int wm_sete3d_from_e(const double *selfe, int64_t np, const double *xp, const double *yp,
                     const double *zp, double zgrid, double xmmin, double ymmin, double zmmin,
                     double dx, double dy, double dz, int64_t nx, int64_t ny, int64_t nz,
                     double *ex, double *ey, double *ez, void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    if (np == 0) return 0;
    WM_REQUIRE(selfe && xp && yp && zp && ex && ey && ez, "wm_sete3d_from_e: NULL array");
    Mesh m = make_mesh(nx, ny, nz, dx, dy, dz, xmmin, ymmin, zmmin, zgrid);
    hipLaunchKernelGGL(k_sete_from_e, dim3(nblk(np)), dim3(256), 0, wm_stream(stream), selfe,
                       (long long)np, xp, yp, zp, m, ex, ey, ez);
    WM_LAUNCHED();
  });
}

int64_t wm_getzmmnt_worksize(int64_t np) {
  return (int64_t)sizeof(double) * NZMMNT * zmmnt_blocks(np);
}

int wm_getzmmnt(int64_t np, const double *xp, const double *yp, const double *zp,
                const double *uxp, const double *uyp, const double *uzp, const double *gaminv,
                int64_t usegaminv, double *mm, void *work, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(np >= 0, "wm_getzmmnt: np < 0");
    WM_REQUIRE(mm && work, "wm_getzmmnt: NULL output or work");
    WM_REQUIRE(np == 0 || (xp && yp && zp && uxp && uyp && uzp && (gaminv || !usegaminv)),
               "wm_getzmmnt: NULL array");
    hipStream_t s = wm_stream(stream);
    int nb = zmmnt_blocks(np);
    double *part = (double *)work;
    hipLaunchKernelGGL(k_zmmnt_partial, dim3(nb), dim3(256), 0, s, (long long)np, xp, yp, zp, uxp,
                       uyp, uzp, gaminv, (int)usegaminv, part);
    WM_LAUNCHED();
    hipLaunchKernelGGL(k_zmmnt_final, dim3(1), dim3(256), 0, s, (const double *)part, nb, mm);
    WM_LAUNCHED();
  });
}

int64_t wm_particle_sort_worksize(int64_t np, int64_t nx, int64_t ny, int64_t nz) {
  return (int64_t)sorted_layout(np, (nx + 1) * (ny + 1) * (nz + 1)).total;
}

int wm_particle_sort(int64_t np, const double *xp, const double *yp, const double *zp,
                     double xmmin, double ymmin, double zmmin, double dx, double dy, double dz,
                     int64_t nx, int64_t ny, int64_t nz, uint32_t *perm, void *work,
                     void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    WM_REQUIRE(np >= 0 && np < (1ll << 31), "wm_particle_sort: np out of range");
    if (np == 0) return 0;
    WM_REQUIRE(xp && yp && zp && perm && work, "wm_particle_sort: NULL array");
    hipStream_t s = wm_stream(stream);
    int64_t nkey = (nx + 1) * (ny + 1) * (nz + 1);
    SortedWork w = sorted_layout(np, nkey);
    char *base = (char *)work;
    unsigned int *keys = (unsigned int *)(base + w.keys), *idx = (unsigned int *)(base + w.idx);
    unsigned int *keys2 = (unsigned int *)(base + w.keys2);
    hipLaunchKernelGGL(k_sortkeys, dim3(nblk(np)), dim3(256), 0, s, (long long)np, xp, yp, zp,
                       xmmin, ymmin, zmmin, dx, dy, dz, (int)nx, (int)ny, (int)nz, keys, idx);
    WM_LAUNCHED();
    // the LSD radix sort is stable: equal keys keep their order, as in the
    // reference's counting sort (top.F:5156-5174)
    unsigned int *sorted_idx = nullptr;
    radix_sort_pairs(keys, idx, keys2, (unsigned int *)perm, np, key_bits(nkey), base + w.temp,
                     s, &sorted_idx);
    if (sorted_idx != (unsigned int *)perm)
      WM_HIP(hipMemcpyAsync(perm, sorted_idx, sizeof(unsigned int) * (size_t)np,
                            hipMemcpyDeviceToDevice, s));
  });
}

int wm_permute_columns(int64_t np, int64_t ncol, const double *const *src, double *const *dst,
                       const uint32_t *perm, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(ncol >= 0 && ncol <= 16, "wm_permute_columns: at most 16 columns");
    if (np == 0 || ncol == 0) return 0;
    WM_REQUIRE(src && dst && perm, "wm_permute_columns: NULL argument");
    Columns c{};
    c.n = (int)ncol;
    for (int i = 0; i < ncol; i++) {
      WM_REQUIRE(src[i] && dst[i] && src[i] != dst[i], "wm_permute_columns: bad column pointer");
      c.src[i] = src[i];
      c.dst[i] = dst[i];
    }
    hipLaunchKernelGGL(k_permute, dim3(nblk(np)), dim3(256), 0, wm_stream(stream), (long long)np, c,
                       (const unsigned int *)perm);
    WM_LAUNCHED();
  });
}

int wm_getgrid3d(int64_t np, const double *x, const double *y, const double *z, double *f,
                 int64_t nx, int64_t ny, int64_t nz, const double *grid, double xmin, double xmax,
                 double ymin, double ymax, double zmin, double zmax, int64_t l2symtry,
                 int64_t l4symtry, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(np >= 0, "wm_getgrid3d: np < 0");
    WM_REQUIRE(nx > 0 && ny > 0 && nz > 0, "wm_getgrid3d: needs nx, ny, nz > 0");
    // getgrid3d's kaboom (dtop.F:2842-2845)
    WM_REQUIRE(xmax != xmin && zmax != zmin, "getgrid3d: the grid extent is zero, xmin==xmax or zmin==zmax");
    if (np == 0) return 0;
    WM_REQUIRE(x && y && z && f && grid, "wm_getgrid3d: NULL array");
    double dxi = nx / (xmax - xmin);
    double dyi = (ymax == ymin) ? 1. : ny / (ymax - ymin);
    double dzi = nz / (zmax - zmin);
    hipLaunchKernelGGL(k_getgrid3d, dim3(nblk(np)), dim3(256), 0, wm_stream(stream),
                       (long long)np, x, y, z, f, (int)nx, (int)ny, (int)nz, grid, xmin, ymin,
                       zmin, dxi, dyi, dzi, (int)(l2symtry != 0), (int)(l4symtry != 0));
    WM_LAUNCHED();
  });
}

int wm_getselfe3d(const double *phi, int64_t nx, int64_t ny, int64_t nz, double *selfe, double dx,
                  double dy, double dz, int64_t lzero, void *stream) {
  WM_ENTRY({
    check_grid(nx, ny, nz);
    WM_REQUIRE(phi && selfe, "wm_getselfe3d: NULL array");
    double tdxi = 0.5 / dx, tdyi = 0.5 / dy, tdzi = 0.5 / dz;
    long long n = (long long)(nx + 1) * (ny + 1) * (nz + 1);
    hipLaunchKernelGGL(k_getselfe, dim3(nblk(n)), dim3(256), 0, wm_stream(stream), phi, (int)nx,
                       (int)ny, (int)nz, selfe, tdxi, tdyi, tdzi, (int)lzero);
    WM_LAUNCHED();
  });
}

}  // extern "C"
