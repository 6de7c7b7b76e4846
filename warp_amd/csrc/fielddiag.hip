// Field diagnostics of the electrostatic step: the line charge
// (gtlchg3dfromrho, pywarp90/dw3d.F:229-320) and the electrostatic energy
// (getese3dfromrhophi, dw3d.F:708-768), as MultiGrid3D calls them on special
// steps (scripts/multigrid.py:433-467).  Every sum keeps the reference's
// order, so the results are the oracle's bits:
//   * a plane's charge is one chain of (nx+1)(ny+1) additions, so each z
//     plane gets a wavefront: the wavefront streams the plane from HBM through
//     registers into LDS while lane 0 adds the previous chunk;
//   * the energy sums rho*phi down each (ix, iy) column first (one thread per
//     column, coalesced across ix), then the columns over iy, then over ix.
// rho has no guard cells; phi has one (nxguardphi = 1).
#include "wm_internal.h"

#include <mutex>

// Device scratch of wm_getese3d_noalloc: the column sums of rho*phi, one
// double per (ix, iy) column, at most ESE_SCRATCH_MAX of them (1 MiB, e.g.
// (nx+1)(ny+1) up to 361 x 361).  Every GPU the library runs on holds its
// own copy for the life of the process.  The host never passes this name to
// a kernel: it passes the address hipGetSymbolAddress returned for the
// current device (ese_scratch_ptr).
constexpr int64_t ESE_SCRATCH_MAX = 1 << 17;
static __device__ __align__(256) double scratch[ESE_SCRATCH_MAX];

namespace {

constexpr int ESE_MAX_DEVICES = 64;

// The scratch's address on the current device, looked up on the first call
// made with that device current and kept.
double *ese_scratch_ptr() {
  static std::mutex mu;
  static double *ptr[ESE_MAX_DEVICES] = {};
  int dev = -1;
  WM_HIP(hipGetDevice(&dev));
  WM_REQUIRE(dev >= 0 && dev < ESE_MAX_DEVICES, "wm_getese3d_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    WM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    WM_REQUIRE(p != nullptr, "wm_getese3d_noalloc: no device address for the scratch");
    ptr[dev] = (double *)p;
  }
  return ptr[dev];
}

constexpr int PS_CHUNK = 2048;             // doubles staged per step (16 KB)
constexpr int PS_PER_LANE = PS_CHUNK / 64;  // registers per lane per chunk

enum { SYM_NONE = 0, SYM_L2 = 1, SYM_L4 = 2 };

// a + q[0] + q[1] + ... + q[L-1], added one after the other.  The next 16
// values are read from LDS while the current 16 are added, so the chain of
// additions does not wait on LDS latency.
__device__ inline double chain_add(double a, const double *q, int L) {
  constexpr int B = 16;
  int j0 = 0;
  if (L >= 2 * B) {
    double cur[B], nxt[B];
#pragma unroll
    for (int j = 0; j < B; j++) cur[j] = q[j];
    for (j0 = B; j0 + B <= L; j0 += B) {
#pragma unroll
      for (int j = 0; j < B; j++) nxt[j] = q[j0 + j];
#pragma unroll
      for (int j = 0; j < B; j++) a = a + cur[j];
#pragma unroll
      for (int j = 0; j < B; j++) cur[j] = nxt[j];
    }
#pragma unroll
    for (int j = 0; j < B; j++) a = a + cur[j];
  }
  for (; j0 < L; j0++) a = a + q[j0];
  return a;
}

// scrtch(iz) of dw3d.F:254-279 for plane blockIdx.x.  MODE SYM_L2: the row
// iy = 0 plus twice the rows iy >= 1; SYM_L4: rho(0,0) + 4*(ix,iy >= 1) +
// 2*(ix >= 1, iy = 0) + 2*(ix = 0, iy >= 1), each partial sum in the SUM
// intrinsic's column-major order.
template <int MODE>
__global__ __launch_bounds__(64) void k_plane_sums(const double *__restrict__ rho, int64_t nx,
                                                   int64_t ny, double *__restrict__ out) {
  __shared__ double buf[2][PS_CHUNK];
  const int lane = threadIdx.x;
  const int64_t npl = (nx + 1) * (ny + 1);
  const double *p = rho + (int64_t)blockIdx.x * npl;
  const int64_t nch = (npl + PS_CHUNK - 1) / PS_CHUNK;
  double r[PS_PER_LANE];
  auto fetch = [&](int64_t c) {
    const int64_t base = c * PS_CHUNK + lane;
#pragma unroll
    for (int j = 0; j < PS_PER_LANE; j++) {
      int64_t k = base + 64 * j;
      r[j] = p[k < npl ? k : npl - 1];
    }
  };
  auto stage = [&](int b) {
#pragma unroll
    for (int j = 0; j < PS_PER_LANE; j++) buf[b][64 * j + lane] = r[j];
  };
  fetch(0);
  stage(0);
  __syncthreads();
  double a = 0., b = 0., c3 = 0.;
  int64_t ix = 0, iy = 0;  // mesh position of the next element (SYM_L4)
  for (int64_t c = 0; c < nch; c++) {
    if (c + 1 < nch) fetch(c + 1);
    if (lane == 0) {
      const double *q = buf[c & 1];
      const int64_t k0 = c * PS_CHUNK;
      const int L = (int)min((int64_t)PS_CHUNK, npl - k0);
      if (MODE == SYM_NONE) {
        a = chain_add(a, q, L);
      } else if (MODE == SYM_L2) {
        // elements k < nx+1 are the row iy = 0
        const int split = (int)max((int64_t)0, min((int64_t)L, nx + 1 - k0));
        a = chain_add(a, q, split);
        b = chain_add(b, q + split, L - split);
      } else {
        for (int j = 0; j < L; j++) {
          const double v = q[j];
          if (iy > 0 && ix > 0)
            a = a + v;
          else if (iy == 0 && ix > 0)
            b = b + v;
          else if (iy > 0)
            c3 = c3 + v;
          if (++ix > nx) {
            ix = 0;
            iy++;
          }
        }
      }
    }
    if (c + 1 < nch) stage((c + 1) & 1);
    __syncthreads();
  }
  if (lane == 0) {
    double s;
    if (MODE == SYM_NONE)
      s = a;
    else if (MODE == SYM_L2)
      s = a + 2. * b;
    else
      s = p[0] + 4. * a + 2. * b + 2. * c3;
    out[blockIdx.x] = s;
  }
}

// dw3d.F:282-316 for one processor: linechg(iz), iz = 0..nzzarr, from the
// plane sums; zz and dzi as the reference computes them on the host side.
__global__ void k_linechg(const double *__restrict__ scr, int64_t nz, int64_t nzzarr, double dzz,
                          double zz, double dzi, double dx, double dy,
                          double *__restrict__ linechg) {
  const int64_t iz = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (iz > nzzarr) return;
  const double zg = ((double)iz * dzz - zz) * dzi;
  const int64_t izg = (int64_t)(zg + (double)nz) - nz;
  const double wzg = zg - (double)izg;
  double v = 0.;
  if (0 <= izg && izg <= nz) v = scr[izg] * dx * dy * (1. - wzg);
  if (0 < izg + 1 && izg + 1 <= nz) v = v + scr[izg + 1] * dx * dy * wzg;
  linechg[iz] = v;
}

// dw3d.F:731-757: scrtch(ix,iy) = sum over iz of rho*phi, then the
// symmetry factor (SYM_L4: 2 on the axes ix = 0 or iy = 0 except the
// corner, 4 elsewhere; SYM_L2: 2 for ix >= 1).
template <int MODE>
__global__ __launch_bounds__(256) void k_ese_columns(const double *__restrict__ rho,
                                                     const double *__restrict__ phi, int64_t nx,
                                                     int64_t ny, int64_t nz,
                                                     double *__restrict__ scr) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t npl = (nx + 1) * (ny + 1);
  if (col >= npl) return;
  const int64_t ix = col % (nx + 1), iy = col / (nx + 1);
  const int64_t pst = (nx + 3) * (ny + 3);
  const double *rp = rho + col;
  const double *pp = phi + (ix + 1) + (nx + 3) * (iy + 1) + pst;
  double s = 0.;
  int64_t iz = 0;
  // batches of U planes; the next batch's loads are issued before the
  // current batch is accumulated
  constexpr int U = 16;
  if (nz + 1 >= 2 * U) {
    double a[U], b[U], na[U], nb[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      a[j] = rp[j * npl];
      b[j] = pp[j * pst];
    }
    for (iz = U; iz + U <= nz + 1; iz += U) {
#pragma unroll
      for (int j = 0; j < U; j++) {
        na[j] = rp[(iz + j) * npl];
        nb[j] = pp[(iz + j) * pst];
      }
#pragma unroll
      for (int j = 0; j < U; j++) s = s + a[j] * b[j];
#pragma unroll
      for (int j = 0; j < U; j++) {
        a[j] = na[j];
        b[j] = nb[j];
      }
    }
#pragma unroll
    for (int j = 0; j < U; j++) s = s + a[j] * b[j];
  }
  for (; iz <= nz; iz++) s = s + rp[iz * npl] * pp[iz * pst];
  if (MODE == SYM_L4) {
    if (ix > 0 && iy > 0)
      s = 4. * s;
    else if (ix > 0 || iy > 0)
      s = 2. * s;
  } else if (MODE == SYM_L2) {
    if (ix > 0) s = 2. * s;
  }
  scr[col] = s;
}

// dw3d.F:759-763: scrtch(ix,0) += scrtch(ix,iy) for iy = 1..ny, per ix.
__global__ void k_ese_rows(double *__restrict__ scr, int64_t nx, int64_t ny) {
  const int64_t ix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ix > nx) return;
  double s = scr[ix];
#pragma unroll 8
  for (int64_t iy = 1; iy <= ny; iy++) s = s + scr[ix + (nx + 1) * iy];
  scr[ix] = s;
}

// dw3d.F:765-769: ese = sum over ix, times dx*dy*dz*0.5.
__global__ void k_ese_final(const double *__restrict__ scr, int64_t nx, double dx, double dy,
                            double dz, double *__restrict__ ese) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double e = 0.;
#pragma unroll 8
  for (int64_t ix = 0; ix <= nx; ix++) e = e + scr[ix];
  ese[0] = dx * dy * dz * 0.5 * e;
}

void check_mesh(const char *who, int64_t nx, int64_t ny, int64_t nz) {
  WM_REQUIRE(nx >= 0 && ny >= 0 && nz >= 0, std::string(who) + ": negative mesh size");
  WM_REQUIRE(nz + 1 <= (1ll << 31) - 1 && (nx + 1) * (ny + 1) < (1ll << 31),
             std::string(who) + ": mesh too large");
}

}  // namespace

extern "C" {

int wm_gtlchg3d(const double *rho, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                double dz, double zgrid, double zmmin, double zzmin, double zbeam, double dzz,
                int64_t nzzarr, int64_t l2symtry, int64_t l4symtry, double *planes,
                double *linechg, void *stream) {
  WM_ENTRY({
    check_mesh("wm_gtlchg3d", nx, ny, nz);
    WM_REQUIRE(nzzarr >= 0 && nzzarr < (1ll << 31), "wm_gtlchg3d: nzzarr out of range");
    WM_REQUIRE(rho && planes && linechg, "wm_gtlchg3d: NULL array");
    hipStream_t s = wm_stream(stream);
    dim3 g((unsigned)(nz + 1)), b(64);
    // l2symtry is tested before l4symtry, as dw3d.F:258-272
    if (l2symtry)
      hipLaunchKernelGGL(k_plane_sums<SYM_L2>, g, b, 0, s, rho, nx, ny, planes);
    else if (l4symtry)
      hipLaunchKernelGGL(k_plane_sums<SYM_L4>, g, b, 0, s, rho, nx, ny, planes);
    else
      hipLaunchKernelGGL(k_plane_sums<SYM_NONE>, g, b, 0, s, rho, nx, ny, planes);
    WM_LAUNCHED();
    const double dzi = (nz == 0) ? 0. : 1. / dz;
    const double zz = zgrid + zmmin - zzmin - zbeam;
    hipLaunchKernelGGL(k_linechg, dim3((unsigned)cdiv(nzzarr + 1, 256)), dim3(256), 0, s,
                       (const double *)planes, nz, nzzarr, dzz, zz, dzi, dx, dy, linechg);
    WM_LAUNCHED();
  });
}

int64_t wm_getese3d_worksize(int64_t nx, int64_t ny) {
  if (nx < 0 || ny < 0) return 0;
  return (int64_t)sizeof(double) * (nx + 1) * (ny + 1);
}

static void ese_launch(const double *rho, const double *phi, int64_t nx, int64_t ny,
                     int64_t nz, double dx, double dy, double dz, int64_t l4symtry,
                     int64_t l2symtry, double *scr, double *ese, hipStream_t s) {
  const int64_t ncol = (nx + 1) * (ny + 1);
  dim3 g((unsigned)cdiv(ncol, 256)), b(256);
  // l4symtry is tested before l2symtry, as dw3d.F:740-757
  if (l4symtry)
    hipLaunchKernelGGL(k_ese_columns<SYM_L4>, g, b, 0, s, rho, phi, nx, ny, nz, scr);
  else if (l2symtry)
    hipLaunchKernelGGL(k_ese_columns<SYM_L2>, g, b, 0, s, rho, phi, nx, ny, nz, scr);
  else
    hipLaunchKernelGGL(k_ese_columns<SYM_NONE>, g, b, 0, s, rho, phi, nx, ny, nz, scr);
  WM_LAUNCHED();
  hipLaunchKernelGGL(k_ese_rows, dim3((unsigned)cdiv(nx + 1, 256)), dim3(256), 0, s, scr, nx,
                     ny);
  WM_LAUNCHED();
  hipLaunchKernelGGL(k_ese_final, dim3(1), dim3(64), 0, s, (const double *)scr, nx, dx, dy, dz,
                     ese);
  WM_LAUNCHED();
}

int wm_getese3d_noalloc(const double *rho, const double *phi, int64_t nx, int64_t ny, int64_t nz,
                        double dx, double dy, double dz, int64_t l4symtry, int64_t l2symtry,
                        double *ese, void *stream) {
  WM_ENTRY({
    check_mesh("wm_getese3d_noalloc", nx, ny, nz);
    WM_REQUIRE(rho && phi && ese, "wm_getese3d_noalloc: NULL array");
    WM_REQUIRE((nx + 1) * (ny + 1) <= ESE_SCRATCH_MAX,
               "wm_getese3d_noalloc: (nx+1)*(ny+1) = " + std::to_string((nx + 1) * (ny + 1)) +
                   " columns exceed the scratch's " + std::to_string(ESE_SCRATCH_MAX) +
                   "; use wm_getese3d with a work buffer");
    // every column sum is written before it is read: nothing to clear
    ese_launch(rho, phi, nx, ny, nz, dx, dy, dz, l4symtry, l2symtry, ese_scratch_ptr(), ese,
               wm_stream(stream));
  });
}

int wm_getese3d(const double *rho, const double *phi, int64_t nx, int64_t ny, int64_t nz,
                double dx, double dy, double dz, int64_t l4symtry, int64_t l2symtry, void *work,
                double *ese, void *stream) {
  WM_ENTRY({
    check_mesh("wm_getese3d", nx, ny, nz);
    WM_REQUIRE(rho && phi && work && ese, "wm_getese3d: NULL array");
    ese_launch(rho, phi, nx, ny, nz, dx, dy, dz, l4symtry, l2symtry, (double *)work, ese,
               wm_stream(stream));
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Streaming probes for scaling the HBM traffic counters: a read-only pass
// over n doubles (the sum is written only when it equals a value that zero
// or finite data never give, so nothing is written) and a write-only pass
// of n doubles, both with 8-byte lanes like the solver's FP64 kernels.
namespace {

__global__ __launch_bounds__(256) void k_probe_read(const double *__restrict__ x,
                                                    double *__restrict__ out, long long n) {
  double v = 0.;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    v += x[i];
  if (v == -1.2345678e300) out[0] = v;
}

__global__ __launch_bounds__(256) void k_probe_write(double *__restrict__ x, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    x[i] = 1.;
}

}  // namespace

extern "C" int wm_stream_probe(double *x, double *out, int64_t n, int64_t mode, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(x && n > 0, "wm_stream_probe: empty array");
    WM_REQUIRE(mode == 0 || mode == 1, "wm_stream_probe: mode must be 0 (read) or 1 (write)");
    hipStream_t s = wm_stream(stream);
    if (mode == 0) {
      WM_REQUIRE(out != nullptr, "wm_stream_probe: NULL out");
      hipLaunchKernelGGL(k_probe_read, dim3(8192), dim3(256), 0, s, (const double *)x, out,
                         (long long)n);
    } else {
      hipLaunchKernelGGL(k_probe_write, dim3(8192), dim3(256), 0, s, x, (long long)n);
    }
    WM_LAUNCHED();
  });
}
