// The full SOR pass (sorpass3d, f3d_mgrid.F:1949-2093: colour 0, then
// colour 1, each with its subgrid-conductor records) with both colours'
// dense updates in one launch, reading phi from one red-black buffer (A)
// and writing the other (B).  Included by mg3d.hip only, after
// mg3d_kernels.h.
//
// Why two buffers: a colour-1 point needs its colour-0 neighbours after the
// colour-0 half sweep.  A block marching up in z can update colour 0 on
// plane z and then colour 1 on plane z-1 (whose colour-0 neighbours on
// planes z-2..z are all new by then) -- but the in-plane neighbours of its
// tile's edge points belong to other blocks, and in place those blocks
// would be overwriting colour 1 while this block still needs the old values.
// With the old phi in A (read only) and the new one going to B, no block
// ever reads a value another block writes in the same launch: a block
// recomputes the colour-0 values of one row beyond each side of its tile
// and one plane beyond each end of its z chunk.  Per node the dense pass
// reads phi and rho once and writes phi once (3 scalars), where two
// in-place half sweeps read phi twice and rho and write phi in two halves
// (4 scalars).
//
// The subgrid-conductor records run as launches of their own around the
// dense one, which only copies their points:
//   * colour-0 records read only old colour-1 values and their own old
//     value, so they are updated before the dense launch, in place in A
//     (record-only k_sor_zmarch on A); the dense launch then reads their
//     new values from A like those of interior (pinned) points and copies
//     every conductor point of both colours to B unchanged;
//   * colour-1 records read new colour-0 values and their own old value,
//     both in B after the dense launch, so they are updated there after it
//     (record-only k_sor_zmarch on B).
// Every record sees the values the reference's two half sweeps give it, so
// the pass is the half sweeps' bit for bit.  The dense launch then has no
// data-dependent loads: every load of plane z+1 is issued, unconditionally
// and at clamped addresses, before plane z is computed, so the one-plane
// prefetch is never drained by a wait for a conditional load.
//
// Tiles: a block owns R consecutive interior rows of every interior element
// of the row (one lane per colour element, 64 * ceil(elements/64) lanes),
// plus a halo row on each side, over a chunk of zc planes.  Phase 1 of step
// z computes colour 0 on plane z for all R+2 rows (halo rows and the chunk's
// halo planes are computed, not stored) into a 3-plane LDS ring; after one
// barrier, phase 2 computes colour 1 on plane z-1 for the R owned rows from
// the ring (x and y neighbours) and the lane's own column (z neighbours,
// kept in registers).
//
// Conditions (checked by the host): six Dirichlet boundaries (the boundary
// conditions between the half sweeps change nothing), no record on an
// interior point, no Neumann subgrid mirror, and B holding the same
// boundary and pinned values as A.
#pragma once

// The loads of one z step, issued one step ahead.
template <typename T>
struct FusedPref {
  T o1p, ym, yp, own, r, eg, r1, hl, hr, sv0, sv1;
  uint32_t mw0, mw1;
};

// One dense pass.  Blocks: by = b % nby (row tile), bz = b / nby (z chunk),
// remapped per XCD.  nl = lanes per row (a multiple of 64), rows = R + 2.
// MASK: points in the conductor mask are copied (their value in A, which
// for colour-0 records is already the new one).  CONV: every unmasked
// point written is compared with save and copied there; each wavefront
// writes its largest |save - new| to cpart[block*waves + wave].
template <typename S, bool CONV, bool MASK>
__global__ __launch_bounds__(1024) void k_sor_fused(
    const S *__restrict__ pa, S *__restrict__ pb, const S *__restrict__ rho, RBGeom g,
    int ixmin, int ixmax, int iymin, int iymax, int izmin, int izmax, typename St<S>::T dxsqic,
    typename St<S>::T dysqic, typename St<S>::T dzsqic, typename St<S>::T spm1,
    const uint32_t *__restrict__ mask, int j0, int nj, int nl, int R, int zc, int nby, int nbt,
    S *__restrict__ save, double *__restrict__ cpart, const int *__restrict__ skip) {
  WM_SKIP();
  typedef typename St<S>::T T;
  extern __shared__ char lds_raw[];
  S *ring = (S *)lds_raw;  // [3][R+2][nj+2]
  const int rowsl = R + 2, rw = nj + 2;
  const int lb = xcd_remap((int)blockIdx.x, nbt);
  const int by = lb % nby, bz = lb / nby;
  const int tid = threadIdx.x;
  const int slot = tid / nl, jl = tid - slot * nl;
  const int lane = tid & 63;
  const bool jin = jl < nj;
  const int j = j0 + (jin ? jl : nj - 1);     // clamped for idle lanes
  const int iy = iymin + by * R - 1 + slot;  // this thread's row (halo rows included)
  const bool rowupd = iy >= iymin && iy <= iymax;
  const bool owned = slot >= 1 && slot <= R && iy <= iymax;
  const int za = izmin + bz * zc;
  const int zb = min(za + zc - 1, izmax);
  const long long pl = (long long)g.W * g.PY;
  const int py = min(iy, g.ny + 1) + 1;  // clamped padded row for idle threads
  const long long rowb = (long long)g.W * py;
  const long long rowc = rowb + j;
  const int jh0 = j0 - 1, jh1 = min(j0 + nj, g.W - 1);  // the row's edge elements
  const S *a0 = pa, *a1 = pa + g.cs;
  S *b0 = pb, *b1 = pb + g.cs;
  const S *r0p = rho, *r1p = rho + g.cs;
  const bool edge0 = lane == 0, edge1 = lane == 63 || jl == nj - 1;
  double cmax = 0.;
  auto ring_at = [&](int z, int s, int e) -> S & {
    return ring[((z % 3 + 3) % 3 * rowsl + s) * rw + e];
  };
  // step z: colour 1 at plane z+1 and at rows y-1, y+1 of plane z, the
  // lane's colour-0 value and source on plane z, the x-neighbour an edge
  // lane takes from memory, the row's two edge elements of colour 0, and on
  // plane z-1 colour 1's source (plus mask words and saved values)
  auto load = [&](int z, FusedPref<T> &p) {
    const long long base = pl * (z + 1);
    p.o1p = St<S>::ld(a1[rowc + base + pl]);
    p.ym = St<S>::ld(a1[rowc + base - g.W]);
    p.yp = St<S>::ld(a1[rowc + base + g.W]);
    p.own = St<S>::ld(a0[rowc + base]);
    p.r = St<S>::ld(r0p[rowc + base]);
    const int p0 = (py + z + 2) & 1;
    p.eg = St<S>::ld(a1[rowb + base + min(max(p0 ? j + 1 : j - 1, 0), g.W - 1)]);
    p.r1 = St<S>::ld(r1p[rowc + base - pl]);
    p.hl = St<S>::ld(a0[rowb + base + jh0]);
    p.hr = St<S>::ld(a0[rowb + base + jh1]);
    if (MASK) {
      p.mw0 = mask[(rowc + base) >> 5];
      p.mw1 = mask[(g.cs + rowc + base - pl) >> 5];
    } else {
      p.mw0 = p.mw1 = 0u;
    }
    if (CONV) {
      p.sv0 = St<S>::ld(save[rowc + base]);
      p.sv1 = St<S>::ld(save[g.cs + rowc + base - pl]);
    } else {
      p.sv0 = p.sv1 = (T)0;
    }
  };
  // column state: colour-1 old at planes z-1, z; colour-0 new at z-2, z-1
  T o1m = St<S>::ld(a1[rowc + pl * (za - 1)]), o1c = St<S>::ld(a1[rowc + pl * za]);
  T c0m2 = 0, c0m1 = 0;
  // one z step from the loads of its plane
  auto step = [&](int z, const FusedPref<T> &c) {
    const int pz = z + 1;
    const long long base = pl * pz;
    // ---- phase 1: colour 0 on plane z
    const int p0 = (py + pz + 1) & 1;
    T sh = p0 ? __shfl_down(o1c, 1, 64) : __shfl_up(o1c, 1, 64);
    if (p0 ? edge1 : edge0) sh = c.eg;
    const T pxm = p0 ? o1c : sh, pxp = p0 ? sh : o1c;
    const int ix = p0 + 2 * j - 1;
    const bool upd = jin && rowupd && z >= izmin && z <= izmax && ix >= ixmin && ix <= ixmax;
    const long long ce0 = rowc + base;
    const bool mk0 = MASK && ((c.mw0 >> (ce0 & 31)) & 1u);
    T c0 = c.own;
    if (upd && !mk0)
      c0 = St<S>::ld(St<S>::st(c.r + (pxm + pxp) * dxsqic + (c.ym + c.yp) * dysqic +
                               (o1m + c.o1p) * dzsqic + spm1 * c.own));
    if (upd && owned && z >= za && z <= zb) {
      b0[ce0] = St<S>::st(c0);
      if (CONV && !mk0) {
        const double d = fabs((double)c.sv0 - (double)c0);
        cmax = d > cmax ? d : cmax;
        save[ce0] = St<S>::st(c0);
      }
    }
    if (jin) ring_at(z, slot, jl + 1) = St<S>::st(c0);
    if (jl == 0) ring_at(z, slot, 0) = St<S>::st(c.hl);
    if (jl == nj - 1) ring_at(z, slot, nj + 1) = St<S>::st(c.hr);
    __syncthreads();
    // ---- phase 2: colour 1 on plane z-1 (owned rows, planes of the chunk)
    const int z2 = z - 1;
    if (owned && jin && z2 >= za && z2 <= zb) {
      const int pz2 = z2 + 1;
      const int q0 = (py + pz2) & 1;  // colour-1 row parity
      const int ix2 = q0 + 2 * j - 1;
      if (ix2 >= ixmin && ix2 <= ixmax) {
        const long long ce1 = rowc + pl * pz2;  // element within the colour-1 array
        const bool mk1 = MASK && ((c.mw1 >> ((g.cs + ce1) & 31)) & 1u);
        T c1 = o1m;
        if (!mk1) {
          const T xl = St<S>::ld(ring_at(z2, slot, jl + (q0 ? 1 : 0)));
          const T xr = St<S>::ld(ring_at(z2, slot, jl + (q0 ? 2 : 1)));
          const T yl = St<S>::ld(ring_at(z2, slot - 1, jl + 1));
          const T yr = St<S>::ld(ring_at(z2, slot + 1, jl + 1));
          c1 = St<S>::ld(St<S>::st(c.r1 + (xl + xr) * dxsqic + (yl + yr) * dysqic +
                                   (c0m2 + c0) * dzsqic + spm1 * o1m));
        }
        b1[ce1] = St<S>::st(c1);
        if (CONV && !mk1) {
          const double d = fabs((double)c.sv1 - (double)c1);
          cmax = d > cmax ? d : cmax;
          save[g.cs + ce1] = St<S>::st(c1);
        }
      }
    }
    c0m2 = c0m1;
    c0m1 = c0;
    o1m = o1c;
    o1c = c.o1p;
  };
  // Sets of loads in turn, so that a step's values are never copied (a
  // copy would wait for them before the next set is issued): two (one step
  // ahead) with the convergence test or narrower types, three (two steps
  // ahead) for the plain FP64 pass; the last steps re-read their own plane
  // into idle sets (unused).
  FusedPref<T> pa_, pb_, pc_;
  if (CONV || sizeof(S) < 8) {
    load(za - 1, pa_);
    for (int z = za - 1; z <= zb + 1; z += 2) {
      load(min(z + 1, zb + 1), pb_);
      step(z, pa_);
      if (z + 1 > zb + 1) break;
      load(min(z + 2, zb + 1), pa_);
      step(z + 1, pb_);
    }
  } else {
    load(za - 1, pa_);
    load(min(za, zb + 1), pb_);
    for (int z = za - 1; z <= zb + 1; z += 3) {
      load(min(z + 2, zb + 1), pc_);
      step(z, pa_);
      if (z + 1 > zb + 1) break;
      load(min(z + 3, zb + 1), pa_);
      step(z + 1, pb_);
      if (z + 2 > zb + 1) break;
      load(min(z + 4, zb + 1), pb_);
      step(z + 2, pc_);
    }
  }
  if (CONV) {
    cmax = wave_max(cmax);
    if (lane == 0) cpart[(long long)blockIdx.x * (blockDim.x >> 6) + (tid >> 6)] = cmax;
  }
}
