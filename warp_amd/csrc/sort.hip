// Stable sort of (key, value) pairs of 32-bit unsigned integers on the GPU:
// a least-significant-digit radix sort, 8-bit digits, one pass per digit up
// to the key's highest bit.  It orders particles by cell for
// particlesortxyzwithcopy (top.F:5095-5203), whose particlesortbyindex is a
// stable counting sort: equal keys keep their original order, which an LSD
// radix sort preserves pass by pass.
//
// One pass over n pairs:
//   1. k_rs_hist: tiles of RS_TILE consecutive pairs count their digits in
//      LDS and write hist[digit * ntiles + tile];
//   2. an exclusive scan of hist (digit-major, so the scan gives every
//      (digit, tile) its first output position: all smaller digits first,
//      then the same digit of the tiles before);
//   3. k_rs_scatter: each tile walks its pairs in order, 256 at a time;
//      within a wavefront the lanes holding the same digit find each other
//      with eight ballots (one per digit bit), the wavefronts' counts per
//      digit are exchanged in LDS, and every pair goes to
//      position(digit, tile) + pairs of that digit before it in the tile.
#include "wm_internal.h"
#include "sort.h"

#include <algorithm>

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
constexpr int RS_BINS = 256;
constexpr int SCAN_BLOCK = 1024;

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint32_t *__restrict__ keys,
                                                        long long n, int shift, int ntiles,
                                                        uint32_t *__restrict__ hist) {
  __shared__ uint32_t cnt[RS_BINS];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ROUNDS; k++) {
    const long long i = base + k * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of n values in three launches: per block of SCAN_BLOCK
// values (in place, block total to sums), over the block totals (one
// block, looping), and the block offsets added back.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *wsum, uint32_t &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
  for (int w = 0; w < nw; w++) {
    if (w < wv) before += wsum[w];
    total += wsum[w];
  }
  __syncthreads();
  return before + incl - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_blocks(uint32_t *__restrict__ a, long long n,
                                                            uint32_t *__restrict__ sums) {
  __shared__ uint32_t wsum[SCAN_BLOCK / 64];
  const long long i = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  const uint32_t v = i < n ? a[i] : 0u;
  uint32_t total;
  const uint32_t ex = block_excl_scan(v, wsum, total);
  if (i < n) a[i] = ex;
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(uint32_t *__restrict__ sums, int nb) {
  __shared__ uint32_t wsum[SCAN_BLOCK / 64];
  uint32_t run = 0;
  for (int c0 = 0; c0 < nb; c0 += SCAN_BLOCK) {
    const int i = c0 + threadIdx.x;
    const uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan(v, wsum, total);
    if (i < nb) sums[i] = run + ex;
    run += total;
  }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(uint32_t *__restrict__ a, long long n,
                                                         const uint32_t *__restrict__ sums) {
  const long long i = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  if (i < n) a[i] += sums[blockIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, long long n, int shift, int ntiles,
    const uint32_t *__restrict__ offs) {
  __shared__ uint32_t base[RS_BINS];
  __shared__ uint32_t wcnt[RS_THREADS / 64][RS_BINS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  base[threadIdx.x] = offs[(long long)threadIdx.x * ntiles + blockIdx.x];
  for (int w = 0; w < RS_THREADS / 64; w++) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  const long long tile0 = (long long)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ROUNDS; k++) {
    const long long i = tile0 + k * RS_THREADS + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    // lanes of this wavefront holding the same digit
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const unsigned long long bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const int before_in_wave = __popcll(peers & below);
    if (valid && before_in_wave == 0) wcnt[wv][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base[d] + (uint32_t)before_in_wave;
      for (int w = 0; w < wv; w++) pos += wcnt[w][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    {
      const int t = threadIdx.x;  // one digit per thread
      uint32_t add = 0;
      for (int w = 0; w < RS_THREADS / 64; w++) {
        add += wcnt[w][t];
        wcnt[w][t] = 0;
      }
      base[t] += add;
    }
    __syncthreads();
  }
}

}  // namespace

size_t radix_sort_worksize(int64_t n) {
  const int64_t ntiles = std::max<int64_t>(1, cdiv(n, RS_TILE));
  const int64_t nh = ntiles * RS_BINS;
  const int64_t nsums = cdiv(nh, SCAN_BLOCK);
  return sizeof(uint32_t) * (size_t)(nh + nsums + 64);
}

uint32_t *radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys2, uint32_t *vals2,
                           int64_t n, int endbit, void *work, hipStream_t s, uint32_t **vals_out) {
  WM_REQUIRE(n >= 0 && n < (1ll << 31), "radix sort: n out of range");
  WM_REQUIRE(endbit >= 1 && endbit <= 32, "radix sort: endbit out of range");
  if (n == 0) {
    *vals_out = vals;
    return keys;
  }
  const int ntiles = (int)std::max<int64_t>(1, cdiv(n, RS_TILE));
  const long long nh = (long long)ntiles * RS_BINS;
  const int nsums = (int)cdiv(nh, SCAN_BLOCK);
  uint32_t *hist = (uint32_t *)work;
  uint32_t *sums = hist + nh;
  uint32_t *ka = keys, *va = vals, *kb = keys2, *vb = vals2;
  for (int shift = 0; shift < endbit; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, (const uint32_t *)ka,
                       (long long)n, shift, ntiles, hist);
    WM_LAUNCHED();
    hipLaunchKernelGGL(k_scan_blocks, dim3(nsums), dim3(SCAN_BLOCK), 0, s, hist, nh, sums);
    WM_LAUNCHED();
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, sums, nsums);
    WM_LAUNCHED();
    hipLaunchKernelGGL(k_scan_add, dim3(nsums), dim3(SCAN_BLOCK), 0, s, hist, nh,
                       (const uint32_t *)sums);
    WM_LAUNCHED();
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, (const uint32_t *)ka,
                       (const uint32_t *)va, kb, vb, (long long)n, shift, ntiles,
                       (const uint32_t *)hist);
    WM_LAUNCHED();
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  *vals_out = va;
  return ka;
}
