// 3-D U-Net kernels for the learned multigrid warm start, gfx950.
//
// Activations are channels-last bf16, x[(z*H + y)*W + x][C], C a multiple of
// 16.  A 3x3x3 convolution (zero padding 1) is an implicit GEMM:
// M = voxels, N = output channels, K = 27 taps x input channels, computed
// with gfx950's v_mfma_f32_16x16x32_bf16 (bf16 operands, fp32
// accumulation).  The block stages its output tile plus a one-voxel halo of
// each 16-channel input block in LDS, so every input value is read from
// memory once per block rather than once per tap.  One MFMA step takes
// K = 32 as two taps of one 16-channel block.  The taps are grouped by
// their z offset dz: the 9 (dy, dx) taps of one dz, numbered t9 = 3*dy +
// dx, form the pairs (2r, 2r+1), r = 0..4, a zero tenth tap padding the
// last, so 15 MFMA steps s = 5*dz + r cover the 27 taps; the weights are
// packed so that each lane's weight operand is 16 contiguous bytes:
//   Wq[((cb*15 + s)*Cout + co)*32 + k],  k < 16: tap (dz, 2r), channel cb*16 + k,
//                                        k >= 16: tap (dz, 2r+1), channel cb*16 + k-16.
// The backward pass over the input is the same convolution with the taps
// mirrored and the channel roles exchanged (a second packing of the same
// weights); the backward pass over the weights accumulates
// dW[co][ci][tap] = sum_v x[v+tap][ci] dy[v][co] with MFMA over voxels, in
// the layout of the fp32 master weights.
#include "wm_internal.h"
#include "../../include/warpmi.h"

#include <algorithm>
#include <type_traits>

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ inline uint16_t f2bf(float f) {
  uint32_t x = __float_as_uint(f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

__device__ inline s4 load4(const uint16_t *p) {
  // 8-byte aligned: channel offsets are multiples of 4
  uint2 v = *(const uint2 *)p;
  s4 r;
  r.x = (short)(v.x & 0xffffu);
  r.y = (short)(v.x >> 16);
  r.z = (short)(v.y & 0xffffu);
  r.w = (short)(v.y >> 16);
  return r;
}

struct Dims {
  int D, H, W;
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));

// Two values to bf16 (round to nearest even) in one instruction
// (v_cvt_pk_bf16_f32), a in the low half.
__device__ inline uint32_t pk_bf16(float a, float b) {
  const f2v v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf2v));
}

// Forward convolution.  A block (4 wavefronts) owns a tile of 16 x VT x 4
// output voxels (x, y, z): wavefront w takes z plane z0 + w and its VT
// rows of 16 voxels, for all NT*16 output channels (VT*NT accumulators of
// 16 voxels x 16 channels).  Per 16-channel input block the block copies
// the tile and its one-voxel halo, 18 x (VT+2) x 6 voxels of 16 channels,
// into LDS (zeros outside the grid), then runs 15 MFMA steps; in step s a
// lane reads its A operand (voxel m of its row, shifted by the tap of step
// s and kg/2, channels 8*(kg%2)..+7) from LDS and its B operands from the
// packed weights, each B operand feeding the VT row tiles.  out = act(conv +
// bias), act the leaky ReLU of slope `slope` (1 = identity); bias may be
// NULL.
template <int NT, int VT>
__global__ __launch_bounds__(256) void k_conv3(const uint16_t *__restrict__ x, int cin,
                                               uint16_t *__restrict__ y, float *__restrict__ yf,
                                               Dims d, const uint16_t *__restrict__ wq,
                                               const float *__restrict__ bias, float slope,
                                               int ntx, int nty) {
  constexpr int HX = 18, HY = VT + 2, HZ = 6, HV = HX * HY * HZ;
  __shared__ __align__(16) uint16_t tile[HV * 16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  int b = blockIdx.x;
  const int tx = b % ntx;
  b /= ntx;
  const int ty = b % nty, tz = b / nty;
  const int x0 = tx * 16, y0 = ty * VT, z0 = tz * 4;
  constexpr int cout = NT * 16;
  const int ncb = cin >> 4;
  f4 acc[VT][NT];
#pragma unroll
  for (int j = 0; j < VT; j++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[j][t] = f4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
  for (int cb = 0; cb < ncb; cb++) {
    __syncthreads();  // every wavefront is done with the previous block's tile
    for (int it = threadIdx.x; it < 2 * HV; it += 256) {
      const int hv = it >> 1, half = it & 1;
      const int hx = hv % HX, r = hv / HX, hy = r % HY, hz = r / HY;
      const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (gx >= 0 && gx < d.W && gy >= 0 && gy < d.H && gz >= 0 && gz < d.D)
        v = *(const uint4 *)(x + (((long long)gz * d.H + gy) * d.W + gx) * cin + cb * 16 +
                             half * 8);
      *(uint4 *)(tile + hv * 16 + half * 8) = v;
    }
    __syncthreads();
    const uint16_t *wcb = wq + (long long)cb * 15 * cout * 32 + 8 * kg;
#pragma unroll 3
    for (int q = 0; q < 15; q++) {
      const int dz = q / 5, t9 = 2 * (q % 5) + (kg >> 1);
      const bool real = t9 < 9;
      const int tt = real ? t9 : 8;
      const int dy = tt / 3, dx = tt % 3;  // shifts + 1
      bf16x8 bq[NT];
#pragma unroll
      for (int t = 0; t < NT; t++)
        bq[t] = *(const bf16x8 *)(wcb + ((long long)q * cout + t * 16 + m) * 32);
#pragma unroll
      for (int j = 0; j < VT; j++) {
        const int hv = ((w + dz) * HY + (j + dy)) * HX + (m + dx);
        bf16x8 a = *(const bf16x8 *)(tile + hv * 16 + 8 * (kg & 1));
        if (!real) a = zero8;
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[j][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq[t], acc[j][t], 0, 0, 0);
      }
    }
  }
  // lane holds D[4*kg + i][m]: voxel x0 + 4*kg + i of row j, channel t*16 + m
  const int zz = z0 + w;
#pragma unroll
  for (int j = 0; j < VT; j++) {
    const int yy = y0 + j;
    if (zz >= d.D || yy >= d.H) continue;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int co = t * 16 + m;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int xx = x0 + 4 * kg + i;
        if (xx >= d.W) continue;
        float v = acc[j][t][i] + bv;
        v = v > 0.f ? v : v * slope;
        const long long vi = ((long long)zz * d.H + yy) * d.W + xx;
        if (y) y[vi * cout + co] = f2bf(v);
        if (yf) yf[vi * cout + co] = v;
      }
    }
  }
}

// Forward convolution marching in z, for 16 output channels and up to 48
// input channels (the finest level of the network) and 32 output channels
// and up to 32 input channels (the next level), NT tiles of 16.  A block (4
// wavefronts) owns a column of 16 x 4VT voxels (x, y) over zc output
// planes; wavefront w takes the rows y0 + w*VT .. +VT-1.  In the packed
// order the 5 steps of one z offset dz use the same (dy, dx) shifts, so one
// LDS read of an input row in halo plane p (16 voxels x 8 channels per
// lane, shifted by (dy, dx)) feeds the MFMAs of the output planes p+1, p
// and p-1 (dz = 0, 1, 2): a third of the LDS reads of a tile that is
// staged per output plane.  Per halo plane the block stages 18 x (4VT+2)
// voxels of all input channels in LDS (zeros outside the grid; a voxel
// takes 2*NCB 16-byte chunks, with no padding: the 16-byte reads then have
// no bank conflicts by SQ_LDS_BANK_CONFLICT, where an odd voxel stride of
// 2*NCB + 1 chunks had 7.5M conflict cycles per launch; two plane buffers
// where two blocks still fit a CU, so that one barrier per plane
// suffices), and the loads of
// the next three planes are in flight during the current plane's MFMAs;
// after plane p, output plane p-1 is complete and is written.  The MFMA
// takes the weights as its A operand (rows: output channels) and the
// voxels as its B operand (columns), so a lane holds 4 consecutive
// channels of one voxel and writes them with one 8-byte (bf16) or 16-byte
// (fp32) store.  The weights stay in registers with one input block
// (NCB = 1), else in LDS (rows of 32 values, the four chunks of each row
// permuted by bits 1-2 of the row index: without the permutation their
// reads had 24M conflict cycles per 48 -> 16 launch, with it none).  Accumulators of three output planes rotate through
// slots 0-2; partial sums of planes outside the chunk fall into slots that
// are cleared before reuse and never written.
template <int VT, int NCB, int NT>
struct Conv3zShape {
  static constexpr int NCH = 2 * NCB, STR = NCH;  // chunks per voxel, voxel stride
  static constexpr int HX = 18, HY = 4 * VT + 2, HV = HX * HY;
  static constexpr int NCHK = HV * NCH;               // chunks staged per plane
  static constexpr int PF = (NCHK + 255) / 256;       // staged chunks per thread
  static constexpr int COUT = 16 * NT;
  static constexpr bool BREG = NCB == 1 && NT == 1;
  static constexpr int BCH = BREG ? 0 : 15 * NCB * COUT * 4;
  // a plane buffer: the plane and a spare chunk; two buffers (one barrier
  // per plane) when two blocks still fit a CU
  static constexpr int PB = HV * STR + 1, SPARE = HV * STR;
  static constexpr bool DB = (2 * PB + BCH) * 16 <= 72 * 1024;
  static constexpr int LDSCH = (DB ? 2 : 1) * PB + BCH;
};

template <int V>
using ic = std::integral_constant<int, V>;

template <int VT, int NCB, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_conv3z(const uint16_t *__restrict__ x, uint16_t *__restrict__ y, float *__restrict__ yf,
              Dims d, const uint16_t *__restrict__ wq, const float *__restrict__ bias,
              float slope, int ntx, int nty, int zc) {
  using C = Conv3zShape<VT, NCB, NT>;
  constexpr int COUT = C::COUT;
  constexpr int CIN = NCB * 16;
  __shared__ __align__(16) uint4 lds[C::LDSCH];
  uint4 *const pl = lds;                                   // staged plane(s)
  uint4 *const bl = lds + (C::DB ? 2 : 1) * C::PB;  // weights (NCB > 1)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m = lane & 15, kg = lane >> 4;
  int b = blockIdx.x;
  const int tx = b % ntx;
  b /= ntx;
  const int ty = b % nty, tzc = b / nty;
  const int x0 = tx * 16, y0 = ty * (4 * VT), z0 = tzc * zc;
  const int z1 = min(z0 + zc, d.D);
  const long long plane = (long long)d.H * d.W * CIN;

  // this thread's staged chunks: offset in a plane of x (clamped into the
  // grid; gin false: the chunk is zero) and in LDS (threads without a k-th
  // chunk store to a spare chunk).  Every load is issued unconditionally, so
  // that waiting for one plane never waits for the planes behind it.
  int goff[C::PF], loff[C::PF];
  bool gin[C::PF];
#pragma unroll
  for (int k = 0; k < C::PF; k++) {
    const int c = tid + 256 * k;
    goff[k] = 0;
    loff[k] = C::SPARE;
    gin[k] = false;
    if (c < C::NCHK) {
      const int hv = c / C::NCH, ch = c - hv * C::NCH;
      const int gx = x0 - 1 + hv % C::HX, gy = y0 - 1 + hv / C::HX;
      loff[k] = hv * C::STR + ch;
      gin[k] = gx >= 0 && gx < d.W && gy >= 0 && gy < d.H;
      if (gin[k]) goff[k] = (gy * d.W + gx) * CIN + ch * 8;
    }
  }
  // planes in flight: halo plane q is loaded into ring entry (q - z0 + 1) % 3,
  // three planes ahead of the one being computed; planes beyond the chunk's
  // last (z1) reload plane z1, planes outside the grid load plane 0 or D-1
  // and are staged as zeros
  uint4 pf[3][C::PF];
  bool pin[3];
  auto fetch = [&](int q, auto e) __attribute__((always_inline)) {
    constexpr int E = decltype(e)::value;
    pin[E] = q >= 0 && q < d.D;
    const int qc = max(0, min(min(q, z1), d.D - 1));
    const uint16_t *xp = x + (long long)qc * plane;
#pragma unroll
    for (int k = 0; k < C::PF; k++) pf[E][k] = *(const uint4 *)(xp + goff[k]);
  };
  // LDS buffer of halo plane q
  auto buf = [&](int q) __attribute__((always_inline)) {
    return C::DB ? pl + ((q - z0 + 1) & 1) * C::PB : pl;
  };
  // ring entry e (halo plane q) into LDS; with one buffer, every wavefront
  // must first be done with the previous plane, with two, with the plane
  // before that, which the barrier after the previous plane's store ensures
  auto stage = [&](auto e, int q) __attribute__((always_inline)) {
    constexpr int E = decltype(e)::value;
    if (!C::DB) __syncthreads();
    uint4 *dst = buf(q);
#pragma unroll
    for (int k = 0; k < C::PF; k++) {
      const uint32_t mk = gin[k] && pin[E] ? 0xffffffffu : 0u;  // values, not addresses
      const uint4 v = pf[E][k];
      dst[loff[k]] = make_uint4(v.x & mk, v.y & mk, v.z & mk, v.w & mk);
    }
    __syncthreads();
  };

  bf16x8 breg[C::BREG ? 15 : 1];
  if constexpr (C::BREG) {
#pragma unroll
    for (int s = 0; s < 15; s++) breg[s] = *(const bf16x8 *)(wq + (s * 16 + m) * 32 + 8 * kg);
  } else {
    const uint4 *wsrc = (const uint4 *)wq;
    for (int i = tid; i < 15 * NCB * COUT * 4; i += 256) {
      const int row = i >> 2;  // (cb * 15 + step) * COUT + output channel
      bl[row * 4 + ((i & 3) ^ ((row >> 1) & 3))] = wsrc[i];
    }
  }
  float bv[NT][4];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int i = 0; i < 4; i++) bv[t][i] = bias ? bias[16 * t + 4 * kg + i] : 0.f;

  // A operand offsets: pair r, this lane's tap t9 = 2r + kg/2 (the tenth,
  // zero tap reads tap 8's voxel and is zeroed), channel half kg%2
  int aoff[5];
#pragma unroll
  for (int r = 0; r < 5; r++) {
    const int t9 = min(2 * r + (kg >> 1), 8);
    aoff[r] = (w * VT + t9 / 3) * C::HX + m + t9 % 3;  // voxel of row j = 0
  }
  const bool fake = (kg >> 1) == 1;
  const bf16x8 zero8 = {};

  f4 acc[3][VT][NT];
#pragma unroll
  for (int q = 0; q < 3; q++)
#pragma unroll
    for (int j = 0; j < VT; j++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[q][j][t] = f4{0.f, 0.f, 0.f, 0.f};

  // dz = 0 adds to the slot of output plane p+1 (S2), dz = 1 to p (S1),
  // dz = 2 to p-1 (S0); DZM selects which of the three run
  auto compute = [&](const uint4 *cur, auto dzm, auto s0, auto s1, auto s2)
                     __attribute__((always_inline)) {
    constexpr int DZM = decltype(dzm)::value, S0 = decltype(s0)::value;
    constexpr int S1 = decltype(s1)::value, S2 = decltype(s2)::value;
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
      // the block's A operands of this input block first, then the MFMAs
      bf16x8 a[5][VT];
#pragma unroll
      for (int r = 0; r < 5; r++)
#pragma unroll
        for (int j = 0; j < VT; j++) {
          a[r][j] = *(const bf16x8 *)&cur[(aoff[r] + j * C::HX) * C::STR + 2 * cb + (kg & 1)];
          if (r == 4 && fake) a[r][j] = zero8;
        }
#pragma unroll
      for (int r = 0; r < 5; r++) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
          bf16x8 bw[3];
#pragma unroll
          for (int dz = 0; dz < 3; dz++) {
            if (!((DZM >> dz) & 1)) continue;
            if constexpr (C::BREG) {
              bw[dz] = breg[dz * 5 + r];
            } else {
              const int row = (cb * 15 + dz * 5 + r) * COUT + 16 * t + m;
              bw[dz] = *(const bf16x8 *)&bl[row * 4 + (kg ^ ((row >> 1) & 3))];
            }
          }
#pragma unroll
          for (int j = 0; j < VT; j++) {
            if (DZM & 1)
              acc[S2][j][t] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[0], a[r][j], acc[S2][j][t], 0, 0, 0);
            if (DZM & 2)
              acc[S1][j][t] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[1], a[r][j], acc[S1][j][t], 0, 0, 0);
            if (DZM & 4)
              acc[S0][j][t] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[2], a[r][j], acc[S0][j][t], 0, 0, 0);
          }
        }
      }
    }
  };
  // lane holds D[4*kg + i][m]: channel 4*kg + i of voxel x0 + m, row j
  auto write = [&](int o, auto s) __attribute__((always_inline)) {
    constexpr int S = decltype(s)::value;
    const int xx = x0 + m;
#pragma unroll
    for (int j = 0; j < VT; j++) {
      const int yy = y0 + w * VT + j;
      if (yy < d.H && xx < d.W) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
          float v[4];
#pragma unroll
          for (int i = 0; i < 4; i++) {
            v[i] = acc[S][j][t][i] + bv[t][i];
            v[i] = v[i] > 0.f ? v[i] : v[i] * slope;
          }
          const long long vi = (((long long)o * d.H + yy) * d.W + xx) * COUT + 16 * t + 4 * kg;
          if (y) *(uint2 *)(y + vi) = make_uint2(pk_bf16(v[0], v[1]), pk_bf16(v[2], v[3]));
          if (yf) *(float4 *)(yf + vi) = make_float4(v[0], v[1], v[2], v[3]);
        }
      }
#pragma unroll
      for (int t = 0; t < NT; t++) acc[S][j][t] = f4{0.f, 0.f, 0.f, 0.f};
    }
  };

  fetch(z0 - 1, ic<0>());
  fetch(z0, ic<1>());
  fetch(z0 + 1, ic<2>());  // z0 + 1 <= z1
  // halo plane z0-1: only its dz = 0 terms (output plane z0, slot 2) count
  stage(ic<0>(), z0 - 1);
  fetch(z0 + 2, ic<0>());
  if (z0 >= 1) compute(buf(z0 - 1), ic<1>(), ic<0>(), ic<1>(), ic<2>());
  // plane z0: no dz = 2 terms (they would belong to plane z0-1)
  stage(ic<1>(), z0);
  fetch(z0 + 3, ic<1>());
  compute(buf(z0), ic<3>(), ic<1>(), ic<2>(), ic<0>());
  int p = z0 + 1;
  // one halo plane p > z0; slots (S0, S1, S2) hold output planes (p-1, p, p+1);
  // the plane is in ring entry S0
  auto step = [&](auto s0, auto s1, auto s2) __attribute__((always_inline)) -> bool {
    stage(s0, p);
    if (p == z1) {  // last halo plane: only its dz = 2 terms (plane z1-1)
      if (p < d.D) compute(buf(p), ic<4>(), s0, s1, s2);
      write(p - 1, s0);
      return true;
    }
    fetch(p + 3, s0);
    compute(buf(p), ic<7>(), s0, s1, s2);
    write(p - 1, s0);
    p++;
    return false;
  };
  while (true) {
    if (step(ic<2>(), ic<0>(), ic<1>())) break;
    if (step(ic<0>(), ic<1>(), ic<2>())) break;
    if (step(ic<1>(), ic<2>(), ic<0>())) break;
  }
}

// Weight gradient in the master layout: dw[(co*cin + ci)*27 + tap] +=
// sum_v x[v+tap][ci] g[v][co].
// One wavefront per (tap, 16 output channels, 16 input channels, chunk of
// voxels); MFMA with M = co, N = ci, K = voxels.
__global__ __launch_bounds__(256) void k_conv3_wgrad(const uint16_t *__restrict__ x, int cin,
                                                     const uint16_t *__restrict__ g, int cout,
                                                     Dims d, float *__restrict__ dw,
                                                     long long nvox, int chunk, int nchunk) {
  int lane = threadIdx.x & 63;
  long long wv = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  int nct = cout >> 4, nci = cin >> 4;
  long long combos = 27ll * nct * nci;
  if (wv >= combos * nchunk) return;
  int ch = (int)(wv / combos);
  int cmb = (int)(wv % combos);
  int tap = cmb / (nct * nci);
  int ct = (cmb / nci) % nct;
  int cit = cmb % nci;
  int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  int m = lane & 15, kg = lane >> 4;
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  long long vbeg = (long long)ch * chunk, vend = std::min<long long>(vbeg + chunk, nvox);
  for (long long vb = vbeg; vb < vend; vb += 16) {
    s4 a, b;
    short av[4], bvv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      long long v = vb + 4 * kg + i;
      uint16_t ga = 0, xb = 0;
      if (v < vend) {
        ga = g[v * cout + ct * 16 + m];
        int xw = (int)(v % d.W);
        long long r = v / d.W;
        int yy = (int)(r % d.H), zz = (int)(r / d.H);
        int sz = zz + dz, sy = yy + dy, sx = xw + dx;
        if (sz >= 0 && sz < d.D && sy >= 0 && sy < d.H && sx >= 0 && sx < d.W)
          xb = x[(((long long)sz * d.H + sy) * d.W + sx) * cin + cit * 16 + m];
      }
      av[i] = (short)ga;
      bvv[i] = (short)xb;
    }
    a = s4{av[0], av[1], av[2], av[3]};
    b = s4{bvv[0], bvv[1], bvv[2], bvv[3]};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  // lane holds D[4*kg + i][m] = dW for co = ct*16 + 4*kg + i, ci = cit*16 + m
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int co = ct * 16 + 4 * kg + i, ci = cit * 16 + m;
    atomicAdd(dw + ((long long)co * cin + ci) * 27 + tap, acc[i]);
  }
}

// Weight gradient, tiled: dw[(co*cin + ci)*27 + tap] += sum_v x[v+tap][ci]
// g[v][co].  A block owns one 16-channel input tile `cit` and the voxels
// [vb, ve) of its chunk; its four wavefronts split the 27 taps (wave w takes
// taps w, w+4, ...) and keep every output tile of their taps in registers
// (NT = cout/16 tiles per tap).  Per step of 16 voxels the block stages in
// LDS, transposed so that each MFMA operand is one 8-byte read, the g tile
// gT[co][16 voxels] and the 27 tap-shifted x tiles xT[tap][16 ci][16 voxels]
// (zero outside the grid); the next step's global loads are issued before
// the current step's MFMAs, and the LDS is double-buffered so one barrier
// per step suffices.  The block's sums go to `part` (no atomics); a second
// kernel adds the chunks in order.
template <int NT>
__global__ __launch_bounds__(256) void k_wgrad_tiles(const uint16_t *__restrict__ x, int cin,
                                                     const uint16_t *__restrict__ g, Dims d,
                                                     long long nvox, long long chunk,
                                                     float *__restrict__ part) {
  constexpr int COUT = NT * 16;
  __shared__ __align__(16) uint16_t gT[2][COUT * 16];
  __shared__ __align__(16) uint16_t xT[2][27 * 16 * 16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const int nci = cin >> 4;
  const int cit = blockIdx.x % nci;
  const long long ch = blockIdx.x / nci;
  const long long vb = ch * chunk, ve = min(vb + chunk, nvox);
  const int nstep = (int)((ve - vb + 15) / 16);
  // staging roles: g: thread t < 4*COUT loads 4 channels (8 bytes) of voxel
  // t / (COUT/4); x: thread t handles voxel i = t & 15 and the (tap, quarter)
  // pairs e = (t >> 4) + 16 k, k = 0..6 (108 pairs)
  const int gi = t / (COUT / 4), gc = (t % (COUT / 4)) * 4;
  const bool gload = t < 4 * COUT;
  const int xi = t & 15;
  uint2 gr = make_uint2(0u, 0u);
  uint2 xr[7];
  auto fetch = [&](int st) {
    long long v0 = vb + 16ll * st;
    long long vg = v0 + gi;
    gr = make_uint2(0u, 0u);
    if (gload && vg < ve) gr = *(const uint2 *)(g + vg * COUT + gc);
    long long vx = v0 + xi;
    int xw = 0, yy = 0, zz = 0;
    bool vin = vx < ve;
    if (vin) {
      xw = (int)(vx % d.W);
      long long r = vx / d.W;
      yy = (int)(r % d.H);
      zz = (int)(r / d.H);
    }
#pragma unroll
    for (int k = 0; k < 7; k++) {
      int e = (t >> 4) + 16 * k;  // 0..111
      xr[k] = make_uint2(0u, 0u);
      if (e < 108 && vin) {
        int tap = e >> 2, q = e & 3;
        int sz = zz + tap / 9 - 1, sy = yy + (tap / 3) % 3 - 1, sx = xw + tap % 3 - 1;
        if (sz >= 0 && sz < d.D && sy >= 0 && sy < d.H && sx >= 0 && sx < d.W)
          xr[k] = *(const uint2 *)(x + (((long long)sz * d.H + sy) * d.W + sx) * cin +
                                   cit * 16 + q * 4);
      }
    }
  };
  auto stage = [&](int b) {
    if (gload) {
      uint16_t *p = gT[b];
      p[(gc + 0) * 16 + gi] = (uint16_t)(gr.x & 0xffffu);
      p[(gc + 1) * 16 + gi] = (uint16_t)(gr.x >> 16);
      p[(gc + 2) * 16 + gi] = (uint16_t)(gr.y & 0xffffu);
      p[(gc + 3) * 16 + gi] = (uint16_t)(gr.y >> 16);
    }
#pragma unroll
    for (int k = 0; k < 7; k++) {
      int e = (t >> 4) + 16 * k;
      if (e < 108) {
        int tap = e >> 2, q = e & 3;
        uint16_t *p = xT[b] + tap * 256 + (q * 4) * 16 + xi;
        p[0] = (uint16_t)(xr[k].x & 0xffffu);
        p[16] = (uint16_t)(xr[k].x >> 16);
        p[32] = (uint16_t)(xr[k].y & 0xffffu);
        p[48] = (uint16_t)(xr[k].y >> 16);
      }
    }
  };
  f4 acc[7][NT];
#pragma unroll
  for (int a = 0; a < 7; a++)
#pragma unroll
    for (int c = 0; c < NT; c++) acc[a][c] = f4{0.f, 0.f, 0.f, 0.f};
  if (nstep > 0) fetch(0);
  for (int st = 0; st < nstep; st++) {
    const int b = st & 1;
    stage(b);
    if (st + 1 < nstep) fetch(st + 1);
    __syncthreads();
    s4 A[NT];
#pragma unroll
    for (int c = 0; c < NT; c++) A[c] = load4(gT[b] + (c * 16 + m) * 16 + 4 * kg);
#pragma unroll
    for (int a = 0; a < 7; a++) {
      int tap = w + 4 * a;
      if (tap < 27) {
        s4 B = load4(xT[b] + tap * 256 + m * 16 + 4 * kg);
#pragma unroll
        for (int c = 0; c < NT; c++)
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[c], B, acc[a][c], 0, 0, 0);
      }
    }
  }
  // lane holds D[4*kg + i][m]: co = c*16 + 4*kg + i, ci = cit*16 + m
  float *pb = part + (long long)blockIdx.x * 27 * COUT * 16;
#pragma unroll
  for (int a = 0; a < 7; a++) {
    int tap = w + 4 * a;
    if (tap >= 27) continue;
#pragma unroll
    for (int c = 0; c < NT; c++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        pb[((long long)tap * COUT + c * 16 + 4 * kg + i) * 16 + m] = acc[a][c][i];
  }
}

// dw[(co*cin + ci)*27 + tap] += sum over the chunks, in chunk order, of the
// partial sums of k_wgrad_tiles.
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float *__restrict__ part, int cin,
                                                      int cout, long long nchunk,
                                                      float *__restrict__ dw) {
  long long o = (long long)blockIdx.x * 256 + threadIdx.x;
  long long n = 27ll * cout * cin;
  if (o >= n) return;
  int tap = (int)(o % 27);
  long long r = o / 27;
  int ci = (int)(r % cin), co = (int)(r / cin);
  int nci = cin >> 4, cit = ci >> 4, cl = ci & 15;
  float s = 0.f;
  for (long long ch = 0; ch < nchunk; ch++)
    s += part[((ch * nci + cit) * 27 + tap) * (long long)cout * 16 + (long long)co * 16 + cl];
  dw[o] = dw[o] + s;
}

// Bias gradient: db[co] += sum_v g[v][co].
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ g, int cout,
                                                   long long nvox, float *__restrict__ db) {
  __shared__ float part[256];
  int co = threadIdx.x % cout;
  int lanes = 256 / cout * cout;
  float s = 0.f;
  if ((int)threadIdx.x < lanes) {
    for (long long v = (long long)blockIdx.x * (256 / cout) + threadIdx.x / cout; v < nvox;
         v += (long long)gridDim.x * (256 / cout))
      s += bf2f(g[v * cout + co]);
  }
  part[threadIdx.x] = (int)threadIdx.x < lanes ? s : 0.f;
  __syncthreads();
  if ((int)threadIdx.x < cout) {
    float t = 0.f;
    for (int k = threadIdx.x; k < lanes; k += cout) t += part[k];
    atomicAdd(db + co, t);
  }
}

// g = g * act'(y): the leaky ReLU keeps the sign, so y > 0 <=> pre > 0.
__global__ __launch_bounds__(256) void k_act_bwd(uint16_t *__restrict__ g,
                                                 const uint16_t *__restrict__ y, float slope,
                                                 long long n) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  if (!(bf2f(y[t]) > 0.f)) g[t] = f2bf(bf2f(g[t]) * slope);
}

// Pack fp32 master weights w[co][ci][tap] into the forward layout, and the
// mirrored/transposed layout of the backward-over-input convolution.
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ w, int cin, int cout,
                                              uint16_t *__restrict__ wq,
                                              uint16_t *__restrict__ wqt) {
  // both layouts hold 30 * cin * cout values (15 steps of two taps)
  const long long n = 30ll * cin * cout;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int k = (int)(t % 32);
  const long long u = t / 32;
  const int half = k >> 4, kl = k & 15;
  {  // forward: u = (cb*15 + s)*cout + co
    const int co = (int)(u % cout);
    const long long v = u / cout;
    const int s = (int)(v % 15), cb = (int)(v / 15);
    const int t9 = 2 * (s % 5) + half, tap = 9 * (s / 5) + t9, ci = cb * 16 + kl;
    wq[t] = t9 < 9 ? f2bf(w[((long long)co * cin + ci) * 27 + tap]) : (uint16_t)0;
  }
  {  // backward over the input: u = (cob*15 + s)*cin + ci, taps mirrored
    const int ci = (int)(u % cin);
    const long long v = u / cin;
    const int s = (int)(v % 15), cob = (int)(v / 15);
    const int t9 = 2 * (s % 5) + half, tap = 9 * (s / 5) + t9, co = cob * 16 + kl;
    wqt[t] = t9 < 9 ? f2bf(w[((long long)co * cin + ci) * 27 + (26 - tap)]) : (uint16_t)0;
  }
}

// Average pooling 2x2x2 (forward) and its adjoint.
__global__ __launch_bounds__(256) void k_pool(const uint16_t *__restrict__ x, Dims d, int c,
                                              uint16_t *__restrict__ y) {
  int D2 = d.D / 2, H2 = d.H / 2, W2 = d.W / 2;
  long long n = (long long)D2 * H2 * W2 * c;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ch = (int)(t % c);
  long long v = t / c;
  int xw = (int)(v % W2);
  long long r = v / W2;
  int yy = (int)(r % H2), zz = (int)(r / H2);
  float s = 0.f;
  for (int k = 0; k < 8; k++) {
    long long sv = ((long long)(2 * zz + (k >> 2)) * d.H + 2 * yy + ((k >> 1) & 1)) * d.W + 2 * xw + (k & 1);
    s += bf2f(x[sv * c + ch]);
  }
  y[t] = f2bf(0.125f * s);
}

__global__ __launch_bounds__(256) void k_pool_bwd(const uint16_t *__restrict__ gy, Dims d,
                                                  int c, uint16_t *__restrict__ gx) {
  long long n = (long long)d.D * d.H * d.W * c;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ch = (int)(t % c);
  long long v = t / c;
  int xw = (int)(v % d.W);
  long long r = v / d.W;
  int yy = (int)(r % d.H), zz = (int)(r / d.H);
  long long cv = ((long long)(zz / 2) * (d.H / 2) + yy / 2) * (d.W / 2) + xw / 2;
  gx[t] = f2bf(0.125f * bf2f(gy[cv * c + ch]));
}

// Nearest upsampling x2 of x (dims d, channels cx) into channels [0, cx) of
// y (dims 2d, cy channels); the adjoint sums the eight children.
__global__ __launch_bounds__(256) void k_up(const uint16_t *__restrict__ x, Dims d, int cx,
                                            uint16_t *__restrict__ y, int cy) {
  long long n = 8ll * d.D * d.H * d.W * cx;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ch = (int)(t % cx);
  long long v = t / cx;
  int W2 = 2 * d.W, H2 = 2 * d.H;
  int xw = (int)(v % W2);
  long long r = v / W2;
  int yy = (int)(r % H2), zz = (int)(r / H2);
  long long sv = ((long long)(zz / 2) * d.H + yy / 2) * d.W + xw / 2;
  y[v * cy + ch] = x[sv * cx + ch];
}

__global__ __launch_bounds__(256) void k_up_bwd(const uint16_t *__restrict__ gy, Dims d,
                                                int cx, int cy, uint16_t *__restrict__ gx) {
  long long n = (long long)d.D * d.H * d.W * cx;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ch = (int)(t % cx);
  long long v = t / cx;
  int xw = (int)(v % d.W);
  long long r = v / d.W;
  int yy = (int)(r % d.H), zz = (int)(r / d.H);
  float s = 0.f;
  for (int k = 0; k < 8; k++) {
    long long cv = ((long long)(2 * zz + (k >> 2)) * (2 * d.H) + 2 * yy + ((k >> 1) & 1)) * (2 * d.W) +
                   2 * xw + (k & 1);
    s += bf2f(gy[cv * cy + ch]);
  }
  gx[t] = f2bf(s);
}

// Channel copy between channels-last tensors: dst[v][doff + c] = src[v][soff + c].
__global__ __launch_bounds__(256) void k_chcopy(const uint16_t *__restrict__ src, int sc, int soff,
                                                uint16_t *__restrict__ dst, int dc, int doff,
                                                int c, long long nvox) {
  long long n = nvox * c;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ch = (int)(t % c);
  long long v = t / c;
  dst[v * dc + doff + ch] = src[v * sc + soff + ch];
}

// The same operations on 8 channels (16 bytes) per thread, for channel
// counts and offsets that are multiples of 8 and fewer than 2^31 groups of
// 8: each channel's arithmetic, and its order, is that of the kernels above.
__device__ inline void unpack8(uint4 u, float *f) {
  f[0] = bf2f((uint16_t)(u.x & 0xffffu)); f[1] = bf2f((uint16_t)(u.x >> 16));
  f[2] = bf2f((uint16_t)(u.y & 0xffffu)); f[3] = bf2f((uint16_t)(u.y >> 16));
  f[4] = bf2f((uint16_t)(u.z & 0xffffu)); f[5] = bf2f((uint16_t)(u.z >> 16));
  f[6] = bf2f((uint16_t)(u.w & 0xffffu)); f[7] = bf2f((uint16_t)(u.w >> 16));
}
__device__ inline uint4 pack8(const float *f) {
  auto two = [](float a, float b) { return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16); };
  return make_uint4(two(f[0], f[1]), two(f[2], f[3]), two(f[4], f[5]), two(f[6], f[7]));
}

__global__ __launch_bounds__(256) void k_pool8(const uint16_t *__restrict__ x, Dims d, int c,
                                               uint16_t *__restrict__ y) {
  const unsigned D2 = d.D / 2, H2 = d.H / 2, W2 = d.W / 2, c8 = c / 8;
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= D2 * H2 * W2 * c8) return;
  const unsigned g = t % c8, v = t / c8;
  const unsigned xw = v % W2, r = v / W2, yy = r % H2, zz = r / H2;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
  for (int k = 0; k < 8; k++) {
    const long long sv =
        ((long long)(2 * zz + (k >> 2)) * d.H + 2 * yy + ((k >> 1) & 1)) * d.W + 2 * xw + (k & 1);
    unpack8(*(const uint4 *)(x + sv * c + 8 * g), f);
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] += f[i];
  }
#pragma unroll
  for (int i = 0; i < 8; i++) s[i] = 0.125f * s[i];
  *(uint4 *)(y + (long long)v * c + 8 * g) = pack8(s);
}

__global__ __launch_bounds__(256) void k_pool_bwd8(const uint16_t *__restrict__ gy, Dims d,
                                                   int c, uint16_t *__restrict__ gx) {
  const unsigned c8 = c / 8, W = d.W, H = d.H;
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (unsigned)d.D * H * W * c8) return;
  const unsigned g = t % c8, v = t / c8;
  const unsigned xw = v % W, r = v / W, yy = r % H, zz = r / H;
  const long long cv = ((long long)(zz / 2) * (H / 2) + yy / 2) * (W / 2) + xw / 2;
  float f[8];
  unpack8(*(const uint4 *)(gy + cv * c + 8 * g), f);
#pragma unroll
  for (int i = 0; i < 8; i++) f[i] = 0.125f * f[i];
  *(uint4 *)(gx + (long long)v * c + 8 * g) = pack8(f);
}

__global__ __launch_bounds__(256) void k_up8(const uint16_t *__restrict__ x, Dims d, int cx,
                                             uint16_t *__restrict__ y, int cy) {
  const unsigned c8 = cx / 8, W2 = 2 * d.W, H2 = 2 * d.H;
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= 8u * d.D * d.H * d.W * c8) return;
  const unsigned g = t % c8, v = t / c8;
  const unsigned xw = v % W2, r = v / W2, yy = r % H2, zz = r / H2;
  const long long sv = ((long long)(zz / 2) * d.H + yy / 2) * d.W + xw / 2;
  *(uint4 *)(y + (long long)v * cy + 8 * g) = *(const uint4 *)(x + sv * cx + 8 * g);
}

// The decoder's input: y (2D, 2H, 2W, cx + cs) = [nearest x2 upsampling of
// x (D, H, W, cx) | skip (2D, 2H, 2W, cs)], each voxel's channels written
// as whole 16-byte groups by consecutive threads.
__global__ __launch_bounds__(256) void k_upcat8(const uint16_t *__restrict__ x, Dims d, int cx,
                                                const uint16_t *__restrict__ skip, int cs,
                                                uint16_t *__restrict__ y) {
  const unsigned cx8 = cx / 8, c8 = (cx + cs) / 8, W2 = 2 * d.W, H2 = 2 * d.H;
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= 8u * d.D * d.H * d.W * c8) return;
  const unsigned g = t % c8, v = t / c8;
  uint4 val;
  if (g < cx8) {
    const unsigned xw = v % W2, r = v / W2, yy = r % H2, zz = r / H2;
    const long long sv = ((long long)(zz / 2) * d.H + yy / 2) * d.W + xw / 2;
    val = *(const uint4 *)(x + sv * cx + 8 * g);
  } else {
    val = *(const uint4 *)(skip + (long long)v * cs + 8 * (g - cx8));
  }
  *(uint4 *)(y + (long long)v * (cx + cs) + 8 * g) = val;
}

__global__ __launch_bounds__(256) void k_up_bwd8(const uint16_t *__restrict__ gy, Dims d,
                                                 int cx, int cy, uint16_t *__restrict__ gx) {
  const unsigned c8 = cx / 8, W = d.W, H = d.H;
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (unsigned)d.D * H * W * c8) return;
  const unsigned g = t % c8, v = t / c8;
  const unsigned xw = v % W, r = v / W, yy = r % H, zz = r / H;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
  for (int k = 0; k < 8; k++) {
    const long long cv = ((long long)(2 * zz + (k >> 2)) * (2 * H) + 2 * yy + ((k >> 1) & 1)) *
                             (2 * W) + 2 * xw + (k & 1);
    unpack8(*(const uint4 *)(gy + cv * cy + 8 * g), f);
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] += f[i];
  }
  *(uint4 *)(gx + (long long)v * cx + 8 * g) = pack8(s);
}

__global__ __launch_bounds__(256) void k_chcopy8(const uint16_t *__restrict__ src, int sc,
                                                 int soff, uint16_t *__restrict__ dst, int dc,
                                                 int doff, int c, unsigned nvox) {
  const unsigned c8 = c / 8;
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= nvox * c8) return;
  const unsigned g = t % c8, v = t / c8;
  *(uint4 *)(dst + (long long)v * dc + doff + 8 * g) =
      *(const uint4 *)(src + (long long)v * sc + soff + 8 * g);
}

__global__ __launch_bounds__(256) void k_act_bwd8(uint16_t *__restrict__ g,
                                                  const uint16_t *__restrict__ y, float slope,
                                                  unsigned n8) {
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= n8) return;
  float fg[8], fy[8];
  uint4 *gp = (uint4 *)g + t;
  const uint4 gu = *gp;
  unpack8(gu, fg);
  unpack8(((const uint4 *)y)[t], fy);
  uint16_t o[8] = {(uint16_t)(gu.x & 0xffffu), (uint16_t)(gu.x >> 16), (uint16_t)(gu.y & 0xffffu),
                   (uint16_t)(gu.y >> 16),     (uint16_t)(gu.z & 0xffffu), (uint16_t)(gu.z >> 16),
                   (uint16_t)(gu.w & 0xffffu), (uint16_t)(gu.w >> 16)};
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (!(fy[i] > 0.f)) o[i] = f2bf(fg[i] * slope);
  *gp = make_uint4((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16),
                   (uint32_t)o[4] | ((uint32_t)o[5] << 16), (uint32_t)o[6] | ((uint32_t)o[7] << 16));
}

__global__ __launch_bounds__(256) void k_add8(uint16_t *__restrict__ a,
                                              const uint16_t *__restrict__ b, unsigned n8) {
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  if (t >= n8) return;
  float fa[8], fb[8];
  unpack8(((const uint4 *)a)[t], fa);
  unpack8(((const uint4 *)b)[t], fb);
#pragma unroll
  for (int i = 0; i < 8; i++) fa[i] = fa[i] + fb[i];
  ((uint4 *)a)[t] = pack8(fa);
}

// a += b elementwise (bf16).
__global__ __launch_bounds__(256) void k_add(uint16_t *__restrict__ a,
                                             const uint16_t *__restrict__ b, long long n) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) a[t] = f2bf(bf2f(a[t]) + bf2f(b[t]));
}

// Adam on fp32 master weights (Kingma & Ba, bias-corrected step).
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, const float *__restrict__ gw,
                                              float *__restrict__ m1, float *__restrict__ m2,
                                              long long n, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  float g = gw[t];
  float a = b1 * m1[t] + (1.f - b1) * g;
  float b = b2 * m2[t] + (1.f - b2) * g * g;
  m1[t] = a;
  m2[t] = b;
  w[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Network input: channels (rho*srho, mask, phi_prev*sphi, 0...) of 16, from
// FP64 natural arrays phi (nz+3, ny+3, nx+3) and rho (nz+1, ny+1, nx+1) over
// the nodes 0..n-1 of each axis, and a 0/1 byte mask of the same nodes.
__global__ __launch_bounds__(256) void k_unet_input(const double *__restrict__ rho,
                                                    const double *__restrict__ phi,
                                                    const uint8_t *__restrict__ mask, int nx,
                                                    int ny, int nz, double srho, double sphi,
                                                    uint16_t *__restrict__ x) {
  long long n = (long long)nx * ny * nz;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  int ix = (int)(t % nx);
  long long r = t / nx;
  int iy = (int)(r % ny), iz = (int)(r / ny);
  double rv = rho[ix + (long long)(nx + 1) * (iy + (long long)(ny + 1) * iz)];
  double pv = phi[(ix + 1) + (long long)(nx + 3) * ((iy + 1) + (long long)(ny + 3) * (iz + 1))];
  uint16_t *o = x + t * 16;
  o[0] = f2bf((float)(rv * srho));
  o[1] = f2bf(mask ? (float)mask[t] : 0.f);
  o[2] = f2bf((float)(pv * sphi));
  for (int k = 3; k < 16; k++) o[k] = 0;
}

// phi(0..n-1) = phi_prev + out/sphi at the nodes that are not pinned
// (mask 0); out is the fp32 network output (channel 0 of 16).
__global__ __launch_bounds__(256) void k_unet_apply(const float *__restrict__ out,
                                                    const uint8_t *__restrict__ mask, int nx,
                                                    int ny, int nz, double sphi,
                                                    double *__restrict__ phi) {
  long long n = (long long)nx * ny * nz;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  if (mask && mask[t]) return;
  int ix = (int)(t % nx);
  long long r = t / nx;
  int iy = (int)(r % ny), iz = (int)(r / ny);
  long long p = (ix + 1) + (long long)(nx + 3) * ((iy + 1) + (long long)(ny + 3) * (iz + 1));
  phi[p] = phi[p] + (double)out[t * 16] / sphi;
}

// Mean-square loss on channel 0 and its gradient (bf16, 16 channels):
// loss += sum (out - tgt)^2 / n, g = 2 (out - tgt) / n.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ out,
                                             const float *__restrict__ tgt, long long n,
                                             uint16_t *__restrict__ g, float *__restrict__ loss) {
  __shared__ float part[256];
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  float e = 0.f;
  if (t < n) {
    float d = out[t * 16] - tgt[t];
    e = d * d / (float)n;
    g[t * 16] = f2bf(2.f * d / (float)n);
    for (int k = 1; k < 16; k++) g[t * 16 + k] = 0;
  }
  part[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// ---------------------------------------------------------------------------
// The learned initial guess works at half resolution: a coarse voxel is the
// 2x2x2 block of mesh nodes (2i..2i+1, 2j..2j+1, 2k..2k+1) of the nodes
// 0..n-1 (n even in every axis), so the (nz/2, ny/2, nx/2) voxel grid is
// divisible by 8 for the network's four levels whenever n is by 16.

// Mean of each 2x2x2 node block of a natural-layout FP64 node array with
// `guard` guard cells (rho: 0, phi: 1), summed in a fixed order.
__global__ __launch_bounds__(256) void k_half_mean(const double *__restrict__ src, int nx, int ny,
                                                   int nz, int guard, float *__restrict__ dst) {
  const int W = nx / 2, H = ny / 2, D = nz / 2;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)W * H * D) return;
  int i = (int)(t % W);
  long long r = t / W;
  int j = (int)(r % H), k = (int)(r / H);
  const long long sx = nx + 1 + 2 * guard, sxy = sx * (ny + 1 + 2 * guard);
  const double *p = src + (2 * i + guard) + sx * (2 * j + guard) + sxy * (2 * k + guard);
  double a = ((p[0] + p[1]) + (p[sx] + p[sx + 1])) + ((p[sxy] + p[sxy + 1]) +
                                                      (p[sxy + sx] + p[sxy + sx + 1]));
  dst[t] = (float)(0.125 * a);
}

// Largest |v| of a block's values into *mx by an integer atomicMax on the
// bit pattern (non-negative floats order as their bits); *mx starts at 0.
__device__ inline void block_absmax(float v, unsigned *mx) {
  __shared__ float part[4];
  v = fabsf(v);
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
    atomicMax(mx, __float_as_uint(m));
  }
}

// d2 = r0 - 2 r1 + r2 (the second time difference of the coarse source) and
// the maxima of |d2| (mx[0]) and |r0| (mx[1]).
__global__ __launch_bounds__(256) void k_second_diff(const float *__restrict__ r0,
                                                     const float *__restrict__ r1,
                                                     const float *__restrict__ r2, long long n,
                                                     float *__restrict__ d2,
                                                     unsigned *__restrict__ mx) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  float a = 0.f, b = 0.f;
  if (t < n) {
    b = r0[t];
    a = (b - 2.f * r1[t]) + r2[t];
    d2[t] = a;
  }
  block_absmax(a, mx);
  __syncthreads();
  block_absmax(b, mx + 1);
}

// Network input (bf16, 16 channels): d2 / max|d2|, the coarse conductor
// fraction, r0 / max|r0|; the other channels zero.
__global__ __launch_bounds__(256) void k_learn_input(const float *__restrict__ d2,
                                                     const float *__restrict__ r0,
                                                     const float *__restrict__ maskh,
                                                     const unsigned *__restrict__ mx,
                                                     long long n, uint16_t *__restrict__ x) {
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  float m0 = __uint_as_float(mx[0]), m1 = __uint_as_float(mx[1]);
  uint16_t *o = x + t * 16;
  o[0] = f2bf(m0 > 0.f ? d2[t] / m0 : 0.f);
  o[1] = f2bf(maskh ? maskh[t] : 0.f);
  o[2] = f2bf(m1 > 0.f ? r0[t] / m1 : 0.f);
  for (int k = 3; k < 16; k++) o[k] = 0;
}

// Training target: the coarse mean of (phi_n - 2 phi_(n-1) + phi_(n-2)),
// the error of the linear extrapolation, over cphys * max|d2|.
__global__ __launch_bounds__(256) void k_learn_target(const double *__restrict__ p0,
                                                      const double *__restrict__ p1,
                                                      const double *__restrict__ p2, int nx,
                                                      int ny, int nz, const unsigned *__restrict__ mx,
                                                      double cphys, float *__restrict__ tgt) {
  const int W = nx / 2, H = ny / 2, D = nz / 2;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)W * H * D) return;
  int i = (int)(t % W);
  long long r = t / W;
  int j = (int)(r % H), k = (int)(r / H);
  const long long sx = nx + 3, sxy = sx * (ny + 3);
  double a = 0.;
  for (int c = 0; c < 8; c++) {
    long long q = (2 * i + (c & 1) + 1) + sx * (2 * j + ((c >> 1) & 1) + 1) +
                  sxy * (2 * k + (c >> 2) + 1);
    a += (p0[q] - 2. * p1[q]) + p2[q];
  }
  double m = (double)__uint_as_float(mx[0]) * cphys;
  tgt[t] = m > 0. ? (float)(0.125 * a / m) : 0.f;
}

// phi(ix,iy,iz) += cphys * max|d2| * (trilinear value of the network's
// channel 0 at the node) on the nodes 1..n-1 that are not pinned (mask 0,
// nodes 0..n-1); coarse voxel centres sit at fine coordinate 2i + 1/2.
__global__ __launch_bounds__(256) void k_learn_apply(const float *__restrict__ out,
                                                     const unsigned *__restrict__ mx,
                                                     double cphys, int nx, int ny, int nz,
                                                     const uint8_t *__restrict__ mask,
                                                     double *__restrict__ phi) {
  const int W = nx / 2, H = ny / 2, D = nz / 2;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)(nx - 1) * (ny - 1) * (nz - 1)) return;
  int ix = 1 + (int)(t % (nx - 1));
  long long r = t / (nx - 1);
  int iy = 1 + (int)(r % (ny - 1)), iz = 1 + (int)(r / (ny - 1));
  if (mask && mask[ix + (long long)nx * (iy + (long long)ny * iz)]) return;
  auto coord = [](int a, int n, int &i0, float &w) {
    float u = 0.5f * ((float)a - 0.5f);
    u = fminf(fmaxf(u, 0.f), (float)(n - 1));
    i0 = min((int)u, n - 2);
    w = u - (float)i0;
  };
  int i0, j0, k0;
  float wx, wy, wz;
  coord(ix, W, i0, wx);
  coord(iy, H, j0, wy);
  coord(iz, D, k0, wz);
  auto v = [&](int a, int b, int c) {
    return out[(((long long)(k0 + c) * H + (j0 + b)) * W + (i0 + a)) * 16];
  };
  float f = (1.f - wz) * ((1.f - wy) * ((1.f - wx) * v(0, 0, 0) + wx * v(1, 0, 0)) +
                          wy * ((1.f - wx) * v(0, 1, 0) + wx * v(1, 1, 0))) +
            wz * ((1.f - wy) * ((1.f - wx) * v(0, 0, 1) + wx * v(1, 0, 1)) +
                  wy * ((1.f - wx) * v(0, 1, 1) + wx * v(1, 1, 1)));
  long long p = (ix + 1) + (long long)(nx + 3) * ((iy + 1) + (long long)(ny + 3) * (iz + 1));
  phi[p] = phi[p] + (double)f * ((double)__uint_as_float(mx[0]) * cphys);
}

inline int nb(long long n) { return (int)std::max<long long>(1, (n + 255) / 256); }
inline bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

}  // namespace

extern "C" {

int wm_unet_conv3(const uint16_t *x, int64_t cin, uint16_t *y, float *yf, int64_t cout,
                  int64_t D, int64_t H, int64_t W, const uint16_t *wp, const float *bias,
                  double slope, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(x && wp && (y || yf), "wm_unet_conv3: NULL argument");
    WM_REQUIRE(cin % 16 == 0 && cin > 0 && cout % 16 == 0 && cout >= 16 && cout <= 128,
               "wm_unet_conv3: channels must be multiples of 16 (cout <= 128)");
    WM_REQUIRE(D > 0 && H > 0 && W > 0, "wm_unet_conv3: empty grid");
    WM_REQUIRE(D * H * W < (1ll << 31), "wm_unet_conv3: grid too large");
    Dims d{(int)D, (int)H, (int)W};
    hipStream_t s = wm_stream(stream);
    float sl = (float)slope;
    if (((cout == 16 && cin <= 48) || (cout == 32 && cin <= 32)) && H * W * cin < (1ll << 31)) {
      // z-marching form; chunks of zc planes, about 2048 blocks or more
      const int ncb = (int)(cin / 16), vt = 2;
      const int ntx = (int)((W + 15) / 16), nty = (int)((H + 4 * vt - 1) / (4 * vt));
      const long long nxy = (long long)ntx * nty;
      const int zc = (int)std::max<long long>(4, std::min<long long>(32, D * nxy / 2048));
      const long long nblk = nxy * ((D + zc - 1) / zc);
      WM_REQUIRE(nblk < (1ll << 31), "wm_unet_conv3: grid too large");
      const dim3 g((unsigned)nblk), bl(256);
      auto run = [&](auto ncbc, auto ntc) {
        hipLaunchKernelGGL((k_conv3z<2, decltype(ncbc)::value, decltype(ntc)::value>), g, bl, 0,
                           s, x, y, yf, d, wp, bias, sl, ntx, nty, zc);
      };
      switch (ncb + 4 * (int)(cout / 16)) {
        case 1 + 4: run(ic<1>(), ic<1>()); break;
        case 2 + 4: run(ic<2>(), ic<1>()); break;
        case 3 + 4: run(ic<3>(), ic<1>()); break;
        case 1 + 8: run(ic<1>(), ic<2>()); break;
        default: run(ic<2>(), ic<2>()); break;
      }
      WM_LAUNCHED();
      return 0;
    }
    // rows of 16 voxels per wavefront: 4 up to 64 output channels, 2 above;
    // 1 where that leaves fewer than 256 blocks (the coarsest levels)
    const int ntx = (int)((W + 15) / 16), nzb = (int)((D + 3) / 4);
    int vt = cout <= 64 ? 4 : 2;
    if ((long long)ntx * ((H + vt - 1) / vt) * nzb < 256) vt = 1;
    const int nty = (int)((H + vt - 1) / vt);
    const long long nblk = (long long)ntx * nty * nzb;
    WM_REQUIRE(nblk < (1ll << 31), "wm_unet_conv3: grid too large");
    const dim3 g((unsigned)nblk), bl(256);
    auto run = [&](auto nt) {
      constexpr int NT = decltype(nt)::value;
      if (vt == 1)
        hipLaunchKernelGGL((k_conv3<NT, 1>), g, bl, 0, s, x, (int)cin, y, yf, d, wp, bias, sl, ntx,
                           nty);
      else
        hipLaunchKernelGGL((k_conv3<NT, NT <= 4 ? 4 : 2>), g, bl, 0, s, x, (int)cin, y, yf, d, wp,
                           bias, sl, ntx, nty);
    };
    switch (cout / 16) {
      case 1: run(ic<1>()); break;
      case 2: run(ic<2>()); break;
      case 3: run(ic<3>()); break;
      case 4: run(ic<4>()); break;
      case 5: run(ic<5>()); break;
      case 6: run(ic<6>()); break;
      case 7: run(ic<7>()); break;
      case 8: run(ic<8>()); break;
      default: throw wm_error("wm_unet_conv3: cout > 128");
    }
    WM_LAUNCHED();
  });
}

int wm_unet_conv3_wgrad(const uint16_t *x, int64_t cin, const uint16_t *g, int64_t cout,
                        int64_t D, int64_t H, int64_t W, float *dw, float *db, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(x && g && dw, "wm_unet_conv3_wgrad: NULL argument");
    WM_REQUIRE(cin % 16 == 0 && cout % 16 == 0 && cin > 0 && cout > 0,
               "wm_unet_conv3_wgrad: channels must be multiples of 16");
    long long nvox = D * H * W;
    Dims d{(int)D, (int)H, (int)W};
    hipStream_t s = wm_stream(stream);
    if (cout <= 64) {
      // about 512 blocks, chunks a multiple of 16 voxels
      long long nci = cin / 16;
      long long nchunk = std::max<long long>(1, std::min<long long>(512 / nci, (nvox + 255) / 256));
      long long chunk = 16 * ((nvox + 16 * nchunk - 1) / (16 * nchunk));
      nchunk = (nvox + chunk - 1) / chunk;
      float *part = nullptr;
      size_t bytes = sizeof(float) * (size_t)(nchunk * nci * 27 * cout * 16);
      WM_HIP(hipMallocAsync((void **)&part, bytes, s));
      unsigned nbk = (unsigned)(nchunk * nci);
      switch (cout / 16) {
        case 1: hipLaunchKernelGGL(k_wgrad_tiles<1>, dim3(nbk), dim3(256), 0, s, x, (int)cin, g, d, nvox, chunk, part); break;
        case 2: hipLaunchKernelGGL(k_wgrad_tiles<2>, dim3(nbk), dim3(256), 0, s, x, (int)cin, g, d, nvox, chunk, part); break;
        case 3: hipLaunchKernelGGL(k_wgrad_tiles<3>, dim3(nbk), dim3(256), 0, s, x, (int)cin, g, d, nvox, chunk, part); break;
        default: hipLaunchKernelGGL(k_wgrad_tiles<4>, dim3(nbk), dim3(256), 0, s, x, (int)cin, g, d, nvox, chunk, part); break;
      }
      WM_LAUNCHED();
      long long n = 27ll * cout * cin;
      hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part,
                         (int)cin, (int)cout, nchunk, dw);
      WM_LAUNCHED();
      WM_HIP(hipFreeAsync(part, s));
    } else {
      int chunk = 4096;
      int nchunk = (int)((nvox + chunk - 1) / chunk);
      long long waves = 27ll * (cout / 16) * (cin / 16) * nchunk;
      hipLaunchKernelGGL(k_conv3_wgrad, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, x,
                         (int)cin, g, (int)cout, d, dw, nvox, chunk, nchunk);
      WM_LAUNCHED();
    }
    if (db) {
      WM_REQUIRE(256 % cout == 0 || cout <= 256, "wm_unet_conv3_wgrad: bad cout");
      hipLaunchKernelGGL(k_bias_grad, dim3(1024), dim3(256), 0, s, g, (int)cout, nvox, db);
      WM_LAUNCHED();
    }
  });
}

int wm_unet_act_bwd(uint16_t *g, const uint16_t *y, double slope, int64_t n, void *stream) {
  WM_ENTRY({
    if (n == 0) return 0;
    if (n % 8 == 0 && n / 8 < (1ll << 31) && al16(g) && al16(y))
      hipLaunchKernelGGL(k_act_bwd8, dim3(nb(n / 8)), dim3(256), 0, wm_stream(stream), g, y,
                         (float)slope, (unsigned)(n / 8));
    else
      hipLaunchKernelGGL(k_act_bwd, dim3(nb(n)), dim3(256), 0, wm_stream(stream), g, y,
                         (float)slope, (long long)n);
    WM_LAUNCHED();
  });
}

int64_t wm_unet_pack_size(int64_t cin, int64_t cout) {
  return cin > 0 && cout > 0 ? 30 * cin * cout : 0;
}

int wm_unet_pack(const float *w, int64_t cin, int64_t cout, uint16_t *wp, uint16_t *wt,
                 void *stream) {
  WM_ENTRY({
    WM_REQUIRE(cin % 16 == 0 && cout % 16 == 0, "wm_unet_pack: channels must be multiples of 16");
    long long n = 30ll * cin * cout;
    hipLaunchKernelGGL(k_pack, dim3(nb(n)), dim3(256), 0, wm_stream(stream), w, (int)cin, (int)cout,
                       wp, wt);
    WM_LAUNCHED();
  });
}

int wm_unet_pool(const uint16_t *x, int64_t D, int64_t H, int64_t W, int64_t c, uint16_t *y,
                 int64_t backward, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(D % 2 == 0 && H % 2 == 0 && W % 2 == 0, "wm_unet_pool: odd dimension");
    Dims d{(int)D, (int)H, (int)W};
    hipStream_t s = wm_stream(stream);
    const bool v8 = c % 8 == 0 && D * H * W * c / 8 < (1ll << 31) && al16(x) && al16(y);
    if (!backward) {
      if (v8)
        hipLaunchKernelGGL(k_pool8, dim3(nb(D * H * W * c / 64)), dim3(256), 0, s, x, d, (int)c, y);
      else
        hipLaunchKernelGGL(k_pool, dim3(nb(D * H * W * c / 8)), dim3(256), 0, s, x, d, (int)c, y);
    } else {
      // x is the gradient at the pooled size, y the gradient at the full size
      if (v8)
        hipLaunchKernelGGL(k_pool_bwd8, dim3(nb(D * H * W * c / 8)), dim3(256), 0, s, x, d, (int)c, y);
      else
        hipLaunchKernelGGL(k_pool_bwd, dim3(nb(D * H * W * c)), dim3(256), 0, s, x, d, (int)c, y);
    }
    WM_LAUNCHED();
  });
}

int wm_unet_up(const uint16_t *x, int64_t D, int64_t H, int64_t W, int64_t cx, uint16_t *y,
               int64_t cy, int64_t backward, void *stream) {
  WM_ENTRY({
    Dims d{(int)D, (int)H, (int)W};
    hipStream_t s = wm_stream(stream);
    const bool v8 = cx % 8 == 0 && cy % 8 == 0 && D * H * W * cx < (1ll << 31) && al16(x) && al16(y);
    if (!backward) {
      if (v8)
        hipLaunchKernelGGL(k_up8, dim3(nb(D * H * W * cx)), dim3(256), 0, s, x, d, (int)cx, y,
                           (int)cy);
      else
        hipLaunchKernelGGL(k_up, dim3(nb(8 * D * H * W * cx)), dim3(256), 0, s, x, d, (int)cx, y,
                           (int)cy);
    } else {
      // x: gradient at the upsampled size (cy channels, first cx used);
      // y: gradient at the coarse size (cx channels)
      if (v8)
        hipLaunchKernelGGL(k_up_bwd8, dim3(nb(D * H * W * cx / 8)), dim3(256), 0, s, x, d, (int)cx,
                           (int)cy, y);
      else
        hipLaunchKernelGGL(k_up_bwd, dim3(nb(D * H * W * cx)), dim3(256), 0, s, x, d, (int)cx,
                           (int)cy, y);
    }
    WM_LAUNCHED();
  });
}

int wm_unet_upcat(const uint16_t *x, int64_t D, int64_t H, int64_t W, int64_t cx,
                  const uint16_t *skip, int64_t cs, uint16_t *y, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(x && skip && y, "wm_unet_upcat: NULL argument");
    Dims d{(int)D, (int)H, (int)W};
    hipStream_t s = wm_stream(stream);
    const long long nf = 8 * D * H * W;  // fine voxels
    if (cx % 8 == 0 && cs % 8 == 0 && nf * (cx + cs) / 8 < (1ll << 31) && al16(x) && al16(skip) &&
        al16(y)) {
      hipLaunchKernelGGL(k_upcat8, dim3(nb(nf * (cx + cs) / 8)), dim3(256), 0, s, x, d, (int)cx,
                         skip, (int)cs, y);
    } else {
      hipLaunchKernelGGL(k_up, dim3(nb(nf * cx)), dim3(256), 0, s, x, d, (int)cx, y,
                         (int)(cx + cs));
      hipLaunchKernelGGL(k_chcopy, dim3(nb(nf * cs)), dim3(256), 0, s, skip, (int)cs, 0, y,
                         (int)(cx + cs), (int)cx, (int)cs, nf);
    }
    WM_LAUNCHED();
  });
}

int wm_unet_chcopy(const uint16_t *src, int64_t sc, int64_t soff, uint16_t *dst, int64_t dc,
                   int64_t doff, int64_t c, int64_t nvox, void *stream) {
  WM_ENTRY({
    if ((sc | soff | dc | doff | c) % 8 == 0 && nvox * c / 8 < (1ll << 31) && al16(src) && al16(dst))
      hipLaunchKernelGGL(k_chcopy8, dim3(nb(nvox * c / 8)), dim3(256), 0, wm_stream(stream), src,
                         (int)sc, (int)soff, dst, (int)dc, (int)doff, (int)c, (unsigned)nvox);
    else
      hipLaunchKernelGGL(k_chcopy, dim3(nb(nvox * c)), dim3(256), 0, wm_stream(stream), src,
                         (int)sc, (int)soff, dst, (int)dc, (int)doff, (int)c, (long long)nvox);
    WM_LAUNCHED();
  });
}

int wm_unet_add(uint16_t *a, const uint16_t *b, int64_t n, void *stream) {
  WM_ENTRY({
    if (n == 0) return 0;
    if (n % 8 == 0 && n / 8 < (1ll << 31) && al16(a) && al16(b))
      hipLaunchKernelGGL(k_add8, dim3(nb(n / 8)), dim3(256), 0, wm_stream(stream), a, b,
                         (unsigned)(n / 8));
    else
      hipLaunchKernelGGL(k_add, dim3(nb(n)), dim3(256), 0, wm_stream(stream), a, b, (long long)n);
    WM_LAUNCHED();
  });
}

int wm_unet_adam(float *w, const float *g, float *m1, float *m2, int64_t n, double lr,
                 double b1, double b2, double eps, int64_t step, void *stream) {
  WM_ENTRY({
    float c1 = 1.f - powf((float)b1, (float)step), c2 = 1.f - powf((float)b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nb(n)), dim3(256), 0, wm_stream(stream), w, g, m1, m2,
                       (long long)n, (float)lr, (float)b1, (float)b2, (float)eps, c1, c2);
    WM_LAUNCHED();
  });
}

int wm_unet_input(const double *rho, const double *phi, const uint8_t *mask, int64_t nx,
                  int64_t ny, int64_t nz, double srho, double sphi, uint16_t *x, void *stream) {
  WM_ENTRY({
    hipLaunchKernelGGL(k_unet_input, dim3(nb(nx * ny * nz)), dim3(256), 0, wm_stream(stream), rho,
                       phi, mask, (int)nx, (int)ny, (int)nz, srho, sphi, x);
    WM_LAUNCHED();
  });
}

int wm_unet_apply(const float *out, const uint8_t *mask, int64_t nx, int64_t ny, int64_t nz,
                  double sphi, double *phi, void *stream) {
  WM_ENTRY({
    hipLaunchKernelGGL(k_unet_apply, dim3(nb(nx * ny * nz)), dim3(256), 0, wm_stream(stream), out,
                       mask, (int)nx, (int)ny, (int)nz, sphi, phi);
    WM_LAUNCHED();
  });
}

int wm_unet_half_mean(const double *src, int64_t nx, int64_t ny, int64_t nz, int64_t guard,
                      float *dst, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(src && dst, "wm_unet_half_mean: NULL array");
    WM_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0,
               "wm_unet_half_mean: nx, ny, nz must be even");
    WM_REQUIRE(guard == 0 || guard == 1, "wm_unet_half_mean: guard must be 0 or 1");
    hipLaunchKernelGGL(k_half_mean, dim3(nb(nx / 2 * (ny / 2) * (nz / 2))), dim3(256), 0,
                       wm_stream(stream), src, (int)nx, (int)ny, (int)nz, (int)guard, dst);
    WM_LAUNCHED();
  });
}

int wm_unet_second_diff(const float *r0, const float *r1, const float *r2, int64_t n, float *d2,
                        uint32_t *mx, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(r0 && r1 && r2 && d2 && mx, "wm_unet_second_diff: NULL array");
    hipLaunchKernelGGL(k_second_diff, dim3(nb(n)), dim3(256), 0, wm_stream(stream), r0, r1, r2,
                       (long long)n, d2, mx);
    WM_LAUNCHED();
  });
}

int wm_unet_learn_input(const float *d2, const float *r0, const float *maskh, const uint32_t *mx,
                        int64_t n, uint16_t *x, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(d2 && r0 && mx && x, "wm_unet_learn_input: NULL array");
    hipLaunchKernelGGL(k_learn_input, dim3(nb(n)), dim3(256), 0, wm_stream(stream), d2, r0,
                       maskh, mx, (long long)n, x);
    WM_LAUNCHED();
  });
}

int wm_unet_learn_target(const double *p0, const double *p1, const double *p2, int64_t nx,
                         int64_t ny, int64_t nz, const uint32_t *mx, double cphys, float *tgt,
                         void *stream) {
  WM_ENTRY({
    WM_REQUIRE(p0 && p1 && p2 && mx && tgt, "wm_unet_learn_target: NULL array");
    WM_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, "wm_unet_learn_target: odd mesh");
    hipLaunchKernelGGL(k_learn_target, dim3(nb(nx / 2 * (ny / 2) * (nz / 2))), dim3(256), 0,
                       wm_stream(stream), p0, p1, p2, (int)nx, (int)ny, (int)nz, mx, cphys, tgt);
    WM_LAUNCHED();
  });
}

int wm_unet_learn_apply(const float *out, const uint32_t *mx, double cphys, int64_t nx,
                        int64_t ny, int64_t nz, const uint8_t *mask, double *phi, void *stream) {
  WM_ENTRY({
    WM_REQUIRE(out && mx && phi, "wm_unet_learn_apply: NULL array");
    WM_REQUIRE(nx >= 4 && ny >= 4 && nz >= 4 && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0,
               "wm_unet_learn_apply: mesh must be even and >= 4");
    hipLaunchKernelGGL(k_learn_apply, dim3(nb((nx - 1) * (ny - 1) * (nz - 1))), dim3(256), 0,
                       wm_stream(stream), out, mx, cphys, (int)nx, (int)ny, (int)nz, mask, phi);
    WM_LAUNCHED();
  });
}

int wm_unet_mse(const float *out, const float *tgt, int64_t n, uint16_t *g, float *loss,
                void *stream) {
  WM_ENTRY({
    hipLaunchKernelGGL(k_mse, dim3(nb(n)), dim3(256), 0, wm_stream(stream), out, tgt,
                       (long long)n, g, loss);
    WM_LAUNCHED();
  });
}

}  // extern "C"
