// The full SOR pass (sorpass3d, f3d_mgrid.F:1949-2093: colour 0, then
// colour 1, each with its subgrid-conductor records) in one launch, reading
// phi from one red-black buffer and writing the other.  Included by
// mg3d.hip only, after mg3d_kernels.h.
//
// Why two buffers: a colour-1 point needs its colour-0 neighbours after the
// colour-0 half sweep.  A block marching up in z can update colour 0 on
// plane z and then colour 1 on plane z-1 (whose colour-0 neighbours on
// planes z-2..z are all new by then) -- but the in-plane neighbours of its
// tile's edge points belong to other blocks, and in place those blocks
// would be overwriting colour 1 while this block still needs the old values.
// With the old phi in buffer A (read only) and the new one going to B, no
// block ever reads a value another block writes: a block recomputes the
// colour-0 values of one row beyond each side of its tile and one plane
// beyond each end of its z chunk.  Per node a pass then reads phi and rho
// once and writes phi once (3 scalars), where two in-place half sweeps read
// phi twice and rho and write phi in two halves (4 scalars), and the two
// launches and the separate record blocks become one launch.
//
// Tiles: a block owns R consecutive interior rows of every interior element
// of the row (one lane per colour element, 64 * ceil(elements/64) lanes),
// plus a halo row on each side, over a chunk of zc planes.  Phase 1 of step
// z computes colour 0 on plane z for all R+2 rows (halo rows and the chunk's
// halo planes are computed, not stored) into a 3-plane LDS ring; after one
// barrier, phase 2 computes colour 1 on plane z-1 for the R owned rows from
// the ring (x and y neighbours) and the lane's own column (z neighbours).
// Points in the conductor mask take the record formula of their colour's
// record list (found through a per-row index) or, when they have no record
// (interior points), keep their value.  The arithmetic is that of
// k_sor_zmarch and its record blocks, term for term, on the same values, so
// the result is the two half sweeps' bit for bit.
//
// Conditions (checked by the host): six Dirichlet boundaries (the boundary
// conditions between the half sweeps change nothing), no record on an
// interior point, no Neumann subgrid mirror, and B holding the same
// boundary values as A.
#pragma once

// Records of one colour for the fused pass: the record list, and per padded
// row (py + PY*pz) the first record of the row (rowptr, PY*PZ+1 entries);
// rj[i] is record i's colour element index within its row.
template <typename T>
struct FusedRecs {
  SubList<T> sub;
  const int *rowptr;
  const int *rj;
};

template <typename T>
__device__ inline int fused_find(const FusedRecs<T> &R, long long row, int j) {
  if (!R.rowptr) return -1;  // no records: a masked point is an interior point
  int a = R.rowptr[row], b = R.rowptr[row + 1];
  for (int i = a; i < b; i++)
    if (R.rj[i] == j) return i;
  return -1;
}

// New value of one point of colour c from its raw neighbours: the dense
// update, or for a masked point its record's form, or (an interior point,
// pinned: kept is set) its old value.  The value is returned as stored.
template <typename S, typename T>
__device__ inline T fused_point(bool masked, const FusedRecs<T> &R, long long row, int j,
                                T r, T own, T pxm, T pxp, T pym, T pyp, T pzm, T pzp, T dxsqic,
                                T dysqic, T dzsqic, T spm1, bool &kept) {
  kept = false;
  if (!masked)
    return St<S>::ld(St<S>::st(r + (pxm + pxp) * dxsqic + (pym + pyp) * dysqic +
                               (pzm + pzp) * dzsqic + spm1 * own));
  int i = fused_find(R, row, j);
  if (i < 0) {
    kept = true;
    return own;
  }
  uint32_t m = R.sub.modes[i];
  const T *k = R.sub.k + 6 * i;
  T xm = side_value<T>(m & 3u, k[0], pxm);
  T xp = side_value<T>((m >> 2) & 3u, k[1], pxp);
  T ym = side_value<T>((m >> 4) & 3u, k[2], pym);
  T yp = side_value<T>((m >> 6) & 3u, k[3], pyp);
  T zm = side_value<T>((m >> 8) & 3u, k[4], pzm);
  T zp = side_value<T>((m >> 10) & 3u, k[5], pzp);
  return St<S>::ld(St<S>::st((r + (xm + xp) * dxsqic + (ym + yp) * dysqic + (zm + zp) * dzsqic) /
                                 R.sub.denom[i] +
                             spm1 * own));
}

// One full pass.  Blocks: by = b % nby (row tile), bz = b / nby (z chunk),
// remapped per XCD.  nl = lanes per row (a multiple of 64), rows = R + 2.
// CONV: every point written is compared with save and copied there; each
// wavefront writes its largest |save - new| to cpart[block*waves + wave].
template <typename S, bool CONV>
__global__ __launch_bounds__(1024) void k_sor_fused(
    const S *__restrict__ pa, S *__restrict__ pb, const S *__restrict__ rho, RBGeom g,
    int ixmin, int ixmax, int iymin, int iymax, int izmin, int izmax, typename St<S>::T dxsqic,
    typename St<S>::T dysqic, typename St<S>::T dzsqic, typename St<S>::T spm1,
    const uint32_t *__restrict__ mask, FusedRecs<typename St<S>::T> rec0,
    FusedRecs<typename St<S>::T> rec1, int j0, int nj, int nl, int R, int zc, int nby, int nbt,
    S *__restrict__ save, double *__restrict__ cpart, const int *__restrict__ skip) {
  WM_SKIP();
  typedef typename St<S>::T T;
  extern __shared__ char lds_raw[];
  S *ring = (S *)lds_raw;  // [3][R+2][nj+2]
  const int rowsl = R + 2, rw = nj + 2;
  const int lb = xcd_remap((int)blockIdx.x, nbt);
  const int by = lb % nby, bz = lb / nby;
  const int tid = threadIdx.x;
  const int slot = tid / nl, jl = tid % nl;
  const int lane = tid & 63;
  const bool jin = jl < nj;
  const int j = j0 + (jin ? jl : nj - 1);      // clamped for idle lanes
  const int iy = iymin + by * R - 1 + slot;   // this thread's row (halo rows included)
  const bool yin = iy <= iymax + 1;            // rows past the tile's end are idle
  const bool owned = slot >= 1 && slot <= R && iy <= iymax;
  const int za = izmin + bz * zc;
  const int zb = min(za + zc - 1, izmax);
  const long long pl = (long long)g.W * g.PY;
  const int py = min(iy, g.ny + 1) + 1;        // clamped padded row for idle threads
  const long long rowc = (long long)g.W * py + j;
  const S *a0 = pa, *a1 = pa + g.cs;
  S *b0 = pb, *b1 = pb + g.cs;
  const S *r0p = rho, *r1p = rho + g.cs;
  double cmax = 0.;
  auto ld = [](const S *p) { return St<S>::ld(*p); };
  auto lds_at = [&](int z, int s, int e) -> S & {
    return ring[((z % 3 + 3) % 3 * rowsl + s) * rw + e];
  };
  // column state: colour-1 old at planes z-1, z, z+1; colour-0 new at z-2,
  // z-1, z (own element, own row)
  T o1m = ld(a1 + rowc + pl * (za - 1)), o1c = ld(a1 + rowc + pl * za);
  T c0m2 = 0, c0m1 = 0;
  // the global loads of step z, issued one step ahead: colour-1 at plane
  // z+1 and at rows y-1, y+1 of plane z, the lane's colour-0 value and
  // source on plane z, the edge lanes' x-neighbour, colour-1's source on
  // plane z-1
  const bool edge0 = lane == 0, edge1 = lane == 63 || jl == nj - 1;
  auto load = [&](int z, T &o1p, T &ym, T &yp, T &own, T &r, T &eg, T &r1) {
    const long long base = pl * (z + 1);
    o1p = ld(a1 + rowc + base + pl);
    ym = ld(a1 + rowc + base - g.W);
    yp = ld(a1 + rowc + base + g.W);
    own = ld(a0 + rowc + base);
    r = ld(r0p + rowc + base);
    const int p0 = (py + z + 2) & 1;
    eg = 0;
    if (p0 ? edge1 : edge0)
      eg = ld(a1 + (long long)g.W * py + base + min(max(p0 ? j + 1 : j - 1, 0), g.W - 1));
    r1 = ld(r1p + rowc + base - pl);
  };
  T n_o1p, n_ym, n_yp, n_own, n_r, n_eg, n_r1;
  load(za - 1, n_o1p, n_ym, n_yp, n_own, n_r, n_eg, n_r1);
  for (int z = za - 1; z <= zb + 1; z++) {
    const int pz = z + 1;
    const long long base = pl * pz;
    const T o1p = n_o1p, ym = n_ym, yp = n_yp, own = n_own, r = n_r, eg = n_eg, r1 = n_r1;
    if (z <= zb) load(z + 1, n_o1p, n_ym, n_yp, n_own, n_r, n_eg, n_r1);
    // ---- phase 1: colour 0 on plane z
    const int p0 = (py + pz + 1) & 1;
    T sh = p0 ? __shfl_down(o1c, 1, 64) : __shfl_up(o1c, 1, 64);
    if (p0 ? edge1 : edge0) sh = eg;
    T pxm = p0 ? o1c : sh, pxp = p0 ? sh : o1c;
    const int ix = p0 + 2 * j - 1;
    const bool upd = jin && yin && iy >= iymin && iy <= iymax && z >= izmin && z <= izmax &&
                     ix >= ixmin && ix <= ixmax;
    T c0 = own;
    if (upd) {
      const long long ce = rowc + base;
      const bool mk = mask && ((mask[ce >> 5] >> (ce & 31)) & 1u);
      bool kept;
      c0 = fused_point<S, T>(mk, rec0, (long long)py + (long long)g.PY * pz, j, r, own, pxm, pxp,
                             ym, yp, o1m, o1p, dxsqic, dysqic, dzsqic, spm1, kept);
      if (owned && z >= za && z <= zb) {
        b0[rowc + base] = St<S>::st(c0);
        if (CONV && !kept) {
          const double d = fabs((double)ld(save + rowc + base) - (double)c0);
          cmax = d > cmax ? d : cmax;
          save[rowc + base] = St<S>::st(c0);
        }
      }
    }
    if (jin) lds_at(z, slot, jl + 1) = St<S>::st(c0);
    // the colour-0 elements beside the lanes (guard or boundary points)
    if (jl == 0) lds_at(z, slot, 0) = a0[(long long)g.W * py + base + (j0 - 1)];
    if (jl == nj - 1) lds_at(z, slot, nj + 1) = a0[(long long)g.W * py + base + min(j0 + nj, g.W - 1)];
    __syncthreads();
    // ---- phase 2: colour 1 on plane z-1 (owned rows, planes of the chunk)
    const int z2 = z - 1;
    if (owned && jin && z2 >= za && z2 <= zb) {
      const int pz2 = z2 + 1;
      const long long base2 = pl * pz2;
      const int q0 = (py + pz2) & 1;  // colour-1 row parity: (1 + py + pz + 1) & 1
      const int ix2 = q0 + 2 * j - 1;
      if (ix2 >= ixmin && ix2 <= ixmax) {
        T xl = St<S>::ld(lds_at(z2, slot, jl + (q0 ? 1 : 0)));
        T xr = St<S>::ld(lds_at(z2, slot, jl + (q0 ? 2 : 1)));
        T yl = St<S>::ld(lds_at(z2, slot - 1, jl + 1));
        T yr = St<S>::ld(lds_at(z2, slot + 1, jl + 1));
        const long long ce = g.cs + rowc + base2;
        const bool mk = mask && ((mask[ce >> 5] >> (ce & 31)) & 1u);
        bool kept;
        T c1 = fused_point<S, T>(mk, rec1, (long long)py + (long long)g.PY * pz2, j, r1, o1m, xl,
                                 xr, yl, yr, c0m2, c0, dxsqic, dysqic, dzsqic, spm1, kept);
        b1[rowc + base2] = St<S>::st(c1);
        if (CONV && !kept) {
          const double d = fabs((double)ld(save + g.cs + rowc + base2) - (double)c1);
          cmax = d > cmax ? d : cmax;
          save[g.cs + rowc + base2] = St<S>::st(c1);
        }
      }
    }
    c0m2 = c0m1;
    c0m1 = c0;
    o1m = o1c;
    o1c = o1p;
  }
  if (CONV) {
    cmax = wave_max(cmax);
    if (lane == 0) cpart[(long long)blockIdx.x * (blockDim.x >> 6) + (tid >> 6)] = cmax;
  }
}
