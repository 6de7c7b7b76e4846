// Absorbing particle boundaries and lost-particle removal on the GPU:
// particleboundaries3d (pywarp90/w3d.F:11292-11450) flags particles that
// left the mesh with gaminv = 0 (xparticleboundaries, top.F:3499-3580, its y
// twin, zparticleboundaries, top.F:3735-3800), and processlostpart ->
// clearpart (top.F:4936-5070, 4730-4893) removes them.
//
// clearpart's fill methods, restated for a parallel machine:
//   1 (the reference's default, top.v clearlostpart = 1): with L live
//     particles, the holes inside [0, L) are filled by the live particles
//     of [L, np), the k-th hole from the front taking the k-th live
//     particle from the back -- exactly what the serial two-pointer loop
//     of top.F:4763-4800 produces.  Done in place: every destination is
//     below L and every source at or above it.
//   2: live particles shifted down, order kept (top.F:4804-4837); done out
//     of place, into a second set of columns.
// Both need, per element, the number of lost particles before it: tiles
// of 1024 particles count theirs (k_lost_tiles, which also applies the
// boundary tests when asked), one block scans the tile counts
// (k_lost_scan), and a third pass ranks elements inside their tile with
// wavefront ballots.  The host reads back only the new count.
#include "wm_internal.h"
#include "../../include/warpmi.h"

#include <algorithm>

namespace {

constexpr int LP_THREADS = 256;
constexpr int LP_PER = 4;
constexpr int LP_TILE = LP_THREADS * LP_PER;
constexpr int LP_MAXCOL = 16;

struct LpCols {
  double *c[LP_MAXCOL];
  int n;
};

struct LpBounds {
  double xmin, xmax, ymin, ymax, zmin, zmax, zgrid;
  int zperiodic;  // pbound0 = pboundnz = periodic: z wraps instead of absorbing
};

// Work layout (int32 words unless noted): tile counts, tile offsets, the
// hole positions and the source positions of method 1, then the totals.
struct LpWork {
  int *tcount, *toff, *holepos, *srcpos;
  long long *tot;  // [0] lost count D, [1] holes below the new end (method 1)
};

static int64_t lp_tiles(int64_t np) { return std::max<int64_t>(1, cdiv(np, LP_TILE)); }

static size_t lp_worksize(int64_t np) {
  int64_t t = lp_tiles(np);
  size_t half = (size_t)(np / 2 + 1);
  return sizeof(long long) * 4 + sizeof(int) * (size_t)(2 * t) + sizeof(int) * 2 * half;
}

static LpWork lp_layout(void *work, int64_t np) {
  LpWork w;
  char *b = (char *)work;
  int64_t t = lp_tiles(np);
  size_t half = (size_t)(np / 2 + 1);
  w.tot = (long long *)b;
  b += sizeof(long long) * 4;
  w.tcount = (int *)b;
  w.toff = w.tcount + t;
  w.holepos = w.toff + t;
  w.srcpos = w.holepos + half;
  return w;
}

// The reference's absorbing tests, in its order (x, y, then z with the
// zgrid shift subtracted from the particle).
__device__ inline bool lp_outside(const LpBounds &b, double x, double y, double z) {
  bool lost = x <= b.xmin || x >= b.xmax;
  lost = lost || y <= b.ymin || y >= b.ymax;
  double zz = z - b.zgrid;
  if (!b.zperiodic) lost = lost || zz >= b.zmax || zz <= b.zmin;
  return lost;
}

// zparticleboundaries with pbound0 = periodic (top.F:3762-3778): a particle
// below the mesh is moved up, one at or above it down, by whole periods.
__device__ inline double lp_zwrap(const LpBounds &b, double z) {
  const double syslen = b.zmax - b.zmin;
  if (z < b.zgrid + b.zmin) z = z + syslen * (double)(long long)((b.zgrid + b.zmax - z) / syslen);
  if (z >= b.zgrid + b.zmax) z = z - syslen * (double)(long long)((z - b.zgrid - b.zmin) / syslen);
  return z;
}

// Exclusive rank of this lane's flag among the block's 256 flags and the
// block's total, from wavefront ballots.
__device__ inline int lp_block_rank(bool f, int *wsum, int &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long b = __ballot(f);
  int r = __popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wv] = __popcll(b);
  __syncthreads();
  int before = 0;
  total = 0;
  for (int k = 0; k < LP_THREADS / 64; k++) {
    int v = wsum[k];
    if (k < wv) before += v;
    total += v;
  }
  __syncthreads();
  return before + r;
}

// Per tile of 1024 particles: optionally flag the ones outside the
// absorbing boundaries (gaminv = 0), then count gaminv <= 0.
template <bool FLAG>
__global__ __launch_bounds__(LP_THREADS) void k_lost_tiles(long long np, const double *__restrict__ x,
                                                           const double *__restrict__ y,
                                                           double *__restrict__ z,
                                                           double *__restrict__ gaminv, LpBounds bd,
                                                           int *__restrict__ tcount) {
  __shared__ int wsum[LP_THREADS / 64];
  const long long base = (long long)blockIdx.x * LP_TILE;
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < LP_PER; k++) {
    long long i = base + k * LP_THREADS + threadIdx.x;
    if (i < np) {
      double g = gaminv[i];
      if (FLAG) {
        double zi = z[i];
        if (bd.zperiodic) {
          double zw = lp_zwrap(bd, zi);
          if (zw != zi) z[i] = zw;
          zi = zw;
        }
        if (lp_outside(bd, x[i], y[i], zi)) {
          g = 0.;
          gaminv[i] = 0.;
        }
      }
      cnt += g <= 0. ? 1 : 0;
    }
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < LP_THREADS / 64; k++) t += wsum[k];
    tcount[blockIdx.x] = t;
  }
}

// One block: exclusive scan of the tile counts into toff, total into tot[0].
__global__ __launch_bounds__(1024) void k_lost_scan(int ntiles, const int *__restrict__ tcount,
                                                    int *__restrict__ toff, long long *tot) {
  __shared__ long long wsum[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  long long run = 0;
  for (int c0 = 0; c0 < ntiles; c0 += 1024) {
    int i = c0 + threadIdx.x;
    long long v = i < ntiles ? tcount[i] : 0;
    long long incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      long long o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    long long before = 0, total = 0;
    for (int k = 0; k < 16; k++) {
      if (k < wv) before += wsum[k];
      total += wsum[k];
    }
    if (i < ntiles) toff[i] = (int)(run + before + incl - v);
    run += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    tot[0] = run;
    tot[1] = 0;
  }
}

// Ranks every element inside its tile.  Method 1: the holes below the new
// end L = np - D record their position by rank, the live particles at or
// above L record theirs by their rank counted from the end; the element at
// L records the number of holes below it.  Method 2: every live particle
// is copied to position (index - lost before it) of the second column set.
template <int METHOD>
__global__ __launch_bounds__(LP_THREADS) void k_lost_rank(long long np, const double *__restrict__ gaminv,
                                                          const int *__restrict__ tcount,
                                                          const int *__restrict__ toff,
                                                          long long *tot, int *__restrict__ holepos,
                                                          int *__restrict__ srcpos, LpCols src,
                                                          LpCols dst) {
  __shared__ int wsum[LP_THREADS / 64];
  const long long D = tot[0];
  if (METHOD == 1 && D == 0) return;
  const long long L = np - D;
  const long long base = (long long)blockIdx.x * LP_TILE;
  const long long tend = std::min<long long>(base + LP_TILE, np);
  const int tc = tcount[blockIdx.x];
  if (METHOD == 1) {
    // nothing to place: no hole below L, or no live particle above it (the
    // tile holding element L always runs: it records the hole count)
    if (tend <= L && tc == 0) return;
    if (base > L && tc == (int)(tend - base)) return;
  } else {
    if (tc == (int)(tend - base)) return;  // every particle of the tile lost
  }
  long long run = toff[blockIdx.x];
  for (int k = 0; k < LP_PER; k++) {
    long long i = base + k * LP_THREADS + threadIdx.x;
    bool in = i < np;
    bool dead = in && gaminv[i] <= 0.;
    int total;
    int r = lp_block_rank(dead, wsum, total);
    long long before = run + r;  // lost particles in [0, i)
    if (METHOD == 1) {
      if (in && i < L && dead) holepos[before] = (int)i;
      if (in && i >= L && !dead) {
        long long after = D - before;                // lost particles in (i, np)
        long long live_after = (np - 1 - i) - after;  // live particles in (i, np)
        srcpos[live_after] = (int)i;
      }
      if (i == L) tot[1] = before;
    } else {
      if (in && !dead) {
        long long d = i - before;
        for (int c = 0; c < src.n; c++) dst.c[c][d] = src.c[c][i];
      }
    }
    run += total;
  }
}

// Method 1: move the k-th live particle from the back into the k-th hole,
// then clear the moved-from slot's uzp and gaminv as clearpart does.
__global__ __launch_bounds__(LP_THREADS) void k_lost_fill(LpCols cols, int iuz, int igam,
                                                          const long long *tot,
                                                          const int *__restrict__ holepos,
                                                          const int *__restrict__ srcpos) {
  const long long m = tot[1];
  for (long long k = (long long)blockIdx.x * LP_THREADS + threadIdx.x; k < m;
       k += (long long)gridDim.x * LP_THREADS) {
    int d = holepos[k], s = srcpos[k];
    for (int c = 0; c < cols.n; c++) cols.c[c][d] = cols.c[c][s];
    cols.c[iuz][s] = 0.;
    cols.c[igam][s] = 0.;
  }
}

__global__ void k_lost_newcount(long long np, const long long *tot, int64_t *nps_out) {
  *nps_out = tot ? np - tot[0] : np;
}

static LpCols lp_cols(int64_t ncol, double *const *cols, const char *who) {
  WM_REQUIRE(cols != nullptr, std::string(who) + ": NULL column list");
  WM_REQUIRE(ncol >= 7 && ncol <= LP_MAXCOL,
             std::string(who) + ": give 7 to 16 columns, xp yp zp uxp uyp uzp gaminv first");
  LpCols c{};
  c.n = (int)ncol;
  for (int i = 0; i < ncol; i++) {
    WM_REQUIRE(cols[i] != nullptr, std::string(who) + ": NULL column");
    c.c[i] = cols[i];
  }
  return c;
}

static void lp_run(int64_t np, LpCols cols, bool flag, const LpBounds &bd, int64_t fillmethod,
                   double *const *dstcols, int64_t *nps_out, void *work, hipStream_t s,
                   const char *who) {
  WM_REQUIRE(np >= 0 && np < (1ll << 31), std::string(who) + ": np out of range");
  WM_REQUIRE(fillmethod == 1 || fillmethod == 2,
             std::string(who) + ": fillmethod must be 1 or 2");
  WM_REQUIRE(nps_out && work, std::string(who) + ": NULL nps_out or work");
  LpCols dst{};
  if (fillmethod == 2) {
    dst = lp_cols(cols.n, dstcols, who);
    for (int i = 0; i < cols.n; i++)
      WM_REQUIRE(dst.c[i] != cols.c[i], std::string(who) + ": fillmethod 2 needs other columns");
  } else {
    WM_REQUIRE(dstcols == nullptr, std::string(who) + ": fillmethod 1 works in place (dst NULL)");
  }
  LpWork w = lp_layout(work, np);
  int nt = (int)lp_tiles(np);
  if (np == 0) {
    hipLaunchKernelGGL(k_lost_newcount, dim3(1), dim3(1), 0, s, 0ll, (const long long *)nullptr,
                       nps_out);
    WM_LAUNCHED();
    return;
  }
  double *x = cols.c[0], *y = cols.c[1], *z = cols.c[2], *g = cols.c[6];
  if (flag)
    hipLaunchKernelGGL(k_lost_tiles<true>, dim3(nt), dim3(LP_THREADS), 0, s, (long long)np, x, y,
                       z, g, bd, w.tcount);
  else
    hipLaunchKernelGGL(k_lost_tiles<false>, dim3(nt), dim3(LP_THREADS), 0, s, (long long)np, x,
                       y, z, g, bd, w.tcount);
  WM_LAUNCHED();
  hipLaunchKernelGGL(k_lost_scan, dim3(1), dim3(1024), 0, s, nt, (const int *)w.tcount, w.toff,
                     w.tot);
  WM_LAUNCHED();
  if (fillmethod == 1) {
    hipLaunchKernelGGL(k_lost_rank<1>, dim3(nt), dim3(LP_THREADS), 0, s, (long long)np,
                       (const double *)g, (const int *)w.tcount, (const int *)w.toff, w.tot,
                       w.holepos, w.srcpos, cols, dst);
    WM_LAUNCHED();
    int nb = (int)std::min<int64_t>(2048, std::max<int64_t>(1, cdiv(np / 2 + 1, LP_THREADS)));
    hipLaunchKernelGGL(k_lost_fill, dim3(nb), dim3(LP_THREADS), 0, s, cols, 5, 6,
                       (const long long *)w.tot, (const int *)w.holepos, (const int *)w.srcpos);
    WM_LAUNCHED();
  } else {
    hipLaunchKernelGGL(k_lost_rank<2>, dim3(nt), dim3(LP_THREADS), 0, s, (long long)np,
                       (const double *)g, (const int *)w.tcount, (const int *)w.toff, w.tot,
                       w.holepos, w.srcpos, cols, dst);
    WM_LAUNCHED();
  }
  hipLaunchKernelGGL(k_lost_newcount, dim3(1), dim3(1), 0, s, (long long)np,
                     (const long long *)w.tot, nps_out);
  WM_LAUNCHED();
}

}  // namespace

extern "C" {

int64_t wm_clearpart_worksize(int64_t np) { return (int64_t)lp_worksize(std::max<int64_t>(np, 0)); }

int wm_particleboundaries3d(int64_t np, int64_t ncol, double *const *cols, double xmmin,
                            double ymmin, double zmmin, double dx, double dy, double dz,
                            int64_t nx, int64_t ny, int64_t nz, double zgrid, int64_t zbound,
                            int64_t fillmethod, double *const *dstcols, int64_t *nps_out,
                            void *work, void *stream) {
  WM_ENTRY({
    LpCols c = lp_cols(ncol, cols, "wm_particleboundaries3d");
    WM_REQUIRE(zbound == 0 || zbound == 2,
               "wm_particleboundaries3d: zbound must be 0 (absorb) or 2 (periodic)");
    // getparticleextent and the calls of particleboundaries3d
    // (w3d.F:11344-11356): the upper extents are xmmin + nx*dx, ...
    LpBounds bd{xmmin, xmmin + nx * dx, ymmin,           ymmin + ny * dy,
                zmmin, zmmin + nz * dz, zgrid, zbound == 2 ? 1 : 0};
    lp_run(np, c, true, bd, fillmethod, dstcols, nps_out, work, wm_stream(stream),
           "wm_particleboundaries3d");
  });
}

int wm_clearpart(int64_t np, int64_t ncol, double *const *cols, int64_t fillmethod,
                 double *const *dstcols, int64_t *nps_out, void *work, void *stream) {
  WM_ENTRY({
    LpCols c = lp_cols(ncol, cols, "wm_clearpart");
    lp_run(np, c, false, LpBounds{}, fillmethod, dstcols, nps_out, work, wm_stream(stream),
           "wm_clearpart");
  });
}

}  // extern "C"
