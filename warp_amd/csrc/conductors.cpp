// Host-side conductor set-up: grid-line intercepts of the supported
// objects, their conversion into interior points and subgrid distances, and
// the per-level clean-up and sort, from pywarp90/f3d_conductors.F and
// pywarp90/f3d_mgrid.F.  This runs once per installed conductor (the
// reference's gridmode = 1); the results are uploaded to the solver by
// wm_mg3d_set_conductors.
#include "wm_internal.h"
#include "../../include/warpmi.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr double LARGEPOS = 1.0e+36;  // pywarp90/top.h:3

// View of one axis' intercept lists: line (a, b) holds n intercepts,
// intercept index fastest, lines ordered with a fastest.
struct Lines {
  double *v;
  double *volt;
  int64_t *cid;
  int64_t n, na;
  double &at(int64_t i, int64_t a, int64_t b) const { return v[(i - 1) + n * (a + na * b)]; }
  double vo(int64_t i, int64_t a, int64_t b) const { return volt[(i - 1) + n * (a + na * b)]; }
  int64_t id(int64_t i, int64_t a, int64_t b) const { return cid[(i - 1) + n * (a + na * b)]; }
};

Lines xlines(const wm_intercepts *c) {
  return Lines{c->xintercepts, c->xvoltages, c->xcondids, c->nxicpt, c->ny + 1};
}
Lines ylines(const wm_intercepts *c) {
  return Lines{c->yintercepts, c->yvoltages, c->ycondids, c->nyicpt, c->nx + 1};
}
Lines zlines(const wm_intercepts *c) {
  return Lines{c->zintercepts, c->zvoltages, c->zcondids, c->nzicpt, c->nx + 1};
}

// Sorted insertion into a line's intercept list (intserticpt,
// f3d_conductors.F:7803-7831).
void insert_icpt(double icpt, int64_t &n, double *line) {
  n = n + 1;
  if (n == 1) {
    line[0] = icpt;
    return;
  }
  for (int64_t i = n - 1; i >= 1; i--) {
    if (icpt >= line[i - 1]) {
      line[i] = icpt;
      break;
    }
    line[i] = line[i - 1];
  }
  if (icpt < line[0]) line[0] = icpt;
}

// Removes two intercepts by shifting I[from..last] = I[from+2..last+2]
// (I is 1-based).  The reference shifts up to n-2 when dropping the pair
// at ic, but only up to n-3 when dropping the pair at ic-1, leaving the
// old value in the new last slot; both are kept as they are.
void drop_pair(double *I, int64_t from, int64_t last, int64_t &n) {
  for (int64_t k = from; k <= last; k++) I[k] = I[k + 2];
  n -= 2;
}

// Fix-up for segments flat in the line's coordinate (ZSrfrvConductorFnew,
// f3d_conductors.F:7318-7397 and 7519-7584): the segment spans lo..hi in
// |transverse| on both sides of cen.
void flat_fix_transverse(double *line, int64_t &n, double lo, double hi, double cen) {
  double *I = line - 1;
  for (int64_t ic = 1; ic < n; ic += 2) {
    if (I[ic] < -hi + cen && I[ic + 1] == -hi + cen) {
      I[ic + 1] = -lo + cen;
    } else if (I[ic] == -lo + cen && I[ic + 1] > -lo + cen) {
      I[ic] = -hi + cen;
      if (ic > 1 && I[ic - 1] == -lo + cen) {
        drop_pair(I, ic - 1, n - 3, n);
        ic -= 2;
      }
    } else if (I[ic] > -hi + cen && I[ic + 1] < -lo + cen) {
      drop_pair(I, ic, n - 2, n);
      ic -= 2;
    } else if (I[ic] < lo + cen && I[ic + 1] == lo + cen) {
      I[ic + 1] = hi + cen;
    } else if (I[ic] == hi + cen && I[ic + 1] > hi + cen) {
      I[ic] = lo + cen;
      if (ic > 1 && I[ic - 1] == hi + cen) {
        drop_pair(I, ic - 1, n - 3, n);
        ic -= 2;
      }
    } else if (I[ic] > lo + cen && I[ic + 1] < hi + cen) {
      drop_pair(I, ic, n - 2, n);
      ic -= 2;
    }
  }
}

// The same for z lines through segments flat in r (7745-7784).
void flat_fix_axial(double *line, int64_t &n, double z1, double z2) {
  double *I = line - 1;
  for (int64_t ic = 1; ic < n; ic += 2) {
    if (I[ic] < z1 && I[ic + 1] == z1) {
      I[ic + 1] = z2;
    } else if (I[ic] == z2 && I[ic + 1] > z2) {
      I[ic] = z1;
      if (ic > 1 && I[ic - 1] == z2) {
        drop_pair(I, ic - 1, n - 3, n);
        ic -= 2;
      }
    } else if (I[ic] > z1 && I[ic + 1] < z2) {
      drop_pair(I, ic, n - 2, n);
      ic -= 2;
    }
  }
}

// Growable conductor lists.
struct SubGrid {
  std::vector<int64_t> indx, numb, ilevel;
  std::vector<double> dels, volt;
  std::vector<int64_t> istart;
  int64_t n() const { return (int64_t)ilevel.size(); }
};

}  // namespace

struct wm_condlist {
  int64_t levels = 0;
  std::vector<int64_t> levelnx, levelny, levelnz;
  std::vector<double> levellx, levelly, levellz;
  std::vector<int64_t> int_indx, int_numb, int_ilevel, int_istart;
  std::vector<double> int_volt;
  SubGrid sub[2];
};

namespace {

// getmglevels (f3d_mgrid.F:3654-3845), serial.
void mglevels(wm_condlist *c, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
              double dz) {
  for (int64_t l = 0; l < c->levels; l++)
    if (c->levellx[l] > 1 || c->levelly[l] > 1 || c->levellz[l] > 1) return;
  c->levelnx.clear();
  c->levelny.clear();
  c->levelnz.clear();
  c->levellx.clear();
  c->levelly.clear();
  c->levellz.clear();
  double lx = 1., ly = 1., lz = 1.;
  for (int64_t level = 0;; level++) {
    c->levelnx.push_back(nx);
    c->levelny.push_back(ny);
    c->levelnz.push_back(nz);
    c->levellx.push_back(lx);
    c->levelly.push_back(ly);
    c->levellz.push_back(lz);
    c->levels = level + 1;
    if (nx < 4 || (ny < 4 && ny > 0) || nz < 4 || level == 100) break;
    int64_t nxc = (nx + 1) / 2, nyc = (ny + 1) / 2, nzc = (nz + 1) / 2;
    if (nxc % 2 == 1) nxc++;
    if (nyc % 2 == 1) nyc++;
    if (nzc % 2 == 1) nzc++;
    double mn = std::min(std::min(dx, dy), dz);
    if (dx > 4. / 3. * mn) nxc = nx;
    if (dy > 4. / 3. * mn) nyc = ny;
    if (dz > 4. / 3. * mn) nzc = nz;
    double dxc = dx * nx / nxc;
    double dyc = nyc == 0 ? dxc : dy * ny / nyc;
    double dzc = dz * nz / nzc;
    lx = lx * (dxc / dx);
    ly = ly * (dyc / dy);
    lz = lz * (dzc / dz);
    nx = nxc;
    ny = nyc;
    nz = nzc;
    dx = dxc;
    dy = dyc;
    dz = dzc;
  }
}

// Per-level clean-up (checkconductors_work, f3d_mgrid.F:4141-4326).
void clean_level(wm_condlist *c, int64_t nx, int64_t ny, int64_t nz, int64_t level) {
  int64_t nmax = std::max(c->sub[0].n(), c->sub[1].n());
  std::vector<int64_t> iii((size_t)((nx + 1) * (ny + 1) * (nz + 1)), 0);
  auto cell = [&](int64_t ix, int64_t iy, int64_t iz) -> int64_t & {
    return iii[(size_t)(ix + (nx + 1) * (iy + (ny + 1) * iz))];
  };
  auto inside = [&](const int64_t *p) {
    return p[0] >= 0 && p[0] <= nx && p[1] >= 0 && p[1] <= ny && p[2] >= 0 && p[2] <= nz;
  };
  for (size_t ic = 0; ic < c->int_ilevel.size(); ic++) {
    if (c->int_ilevel[ic] != level) continue;
    const int64_t *p = &c->int_indx[3 * ic];
    if (!inside(p) || cell(p[0], p[1], p[2]) == nmax + 1) {
      c->int_ilevel[ic] = -1;
      continue;
    }
    cell(p[0], p[1], p[2]) = nmax + 1;
  }
  for (int par = 0; par < 2; par++) {
    SubGrid &s = c->sub[par];
    for (int64_t ic = 0; ic < s.n(); ic++) {
      if (s.ilevel[ic] != level) continue;
      const int64_t *p = &s.indx[3 * ic];
      if (!inside(p) || cell(p[0], p[1], p[2]) == nmax + 1) {
        s.ilevel[ic] = -1;
        continue;
      }
      int64_t &slot = cell(p[0], p[1], p[2]);
      if (slot == 0) {
        slot = ic + 1;
        continue;
      }
      int64_t i = slot - 1;
      if (s.ilevel[ic] != s.ilevel[i]) {
        slot = ic + 1;
        continue;
      }
      // Merge into the earlier point: keep, per direction, the nearer
      // surface.
      for (int id = 0; id < 6; id++) {
        double dn = s.dels[6 * ic + id], de = s.dels[6 * i + id];
        if (std::fabs(dn) < 1. && std::fabs(dn) < std::fabs(de)) {
          s.dels[6 * i + id] = dn;
          s.volt[6 * i + id] = s.volt[6 * ic + id];
          s.numb[6 * i + id] = s.numb[6 * ic + id];
        }
      }
      s.ilevel[ic] = -1;
    }
  }
}

// Stable partition by level (isortconductor + swaps, f3d_mgrid.F:4078-4139).
template <typename T>
void gather_rows(std::vector<T> &a, const std::vector<int64_t> &order, int width) {
  std::vector<T> b(order.size() * width);
  for (size_t i = 0; i < order.size(); i++)
    for (int k = 0; k < width; k++) b[i * width + k] = a[order[i] * width + k];
  a.swap(b);
}

std::vector<int64_t> level_order(const std::vector<int64_t> &ilevel, int64_t levels,
                                 std::vector<int64_t> &istart) {
  std::vector<int64_t> order;
  istart.assign(levels + 1, 0);
  for (int64_t l = 0; l < levels; l++) {
    istart[l] = (int64_t)order.size();
    for (size_t i = 0; i < ilevel.size(); i++)
      if (ilevel[i] == l) order.push_back((int64_t)i);
  }
  istart[levels] = (int64_t)order.size();
  return order;
}

}  // namespace

extern "C" {

int wm_zcylinder_intercepts(double rad, double length, double xcent, double ycent, double zcent,
                            wm_intercepts *icpt, double fuzz, int64_t outside) {
  WM_ENTRY({
    WM_REQUIRE(icpt != nullptr, "wm_zcylinder_intercepts: NULL intercepts");
    int64_t want = outside ? 4 : 2;
    WM_REQUIRE(icpt->nxicpt == want && icpt->nyicpt == want && icpt->nzicpt == 2,
               "wm_zcylinder_intercepts: intercept array sizes do not match the object");
    Lines X = xlines(icpt), Y = ylines(icpt), Z = zlines(icpt);
    double zlo = -length * 0.5, zhi = +length * 0.5;
    double r2 = rad * rad;
    // Lines along x (y varies within a plane) and along y: the chord of the
    // circle of radius rad at the transverse offset t.
    auto transverse = [&](const Lines &L, int64_t na, int64_t nb, double amin, double da,
                          double acent, double bmin, double db, double cent) {
      for (int64_t b = 0; b <= nb; b++) {
        double zz = bmin + b * db - zcent;
        bool inz = zlo - fuzz <= zz && zz <= zhi + fuzz;
        for (int64_t a = 0; a <= na; a++) {
          double t = amin + a * da - acent;
          double h = std::sqrt(std::max(0., r2 - t * t));
          if (!outside) {
            if (inz && std::fabs(t) <= rad) {
              L.at(1, a, b) = cent - h;
              L.at(2, a, b) = cent + h;
            } else {
              L.at(1, a, b) = LARGEPOS;
              L.at(2, a, b) = LARGEPOS;
            }
          } else if (inz) {
            L.at(1, a, b) = -LARGEPOS;
            if (std::fabs(t) <= rad) {
              L.at(2, a, b) = cent - h;
              L.at(3, a, b) = cent + h;
              L.at(4, a, b) = +LARGEPOS;
            } else {
              for (int k = 2; k <= 4; k++) L.at(k, a, b) = +LARGEPOS;
            }
          } else {
            for (int k = 1; k <= 4; k++) L.at(k, a, b) = +LARGEPOS;
          }
        }
      }
    };
    transverse(X, icpt->ny, icpt->nz, icpt->ymmin, icpt->dy, ycent, icpt->zmmin, icpt->dz, xcent);
    transverse(Y, icpt->nx, icpt->nz, icpt->xmmin, icpt->dx, xcent, icpt->zmmin, icpt->dz, ycent);
    for (int64_t iy = 0; iy <= icpt->ny; iy++) {
      double yy = icpt->ymmin + iy * icpt->dy - ycent;
      for (int64_t ix = 0; ix <= icpt->nx; ix++) {
        double xx = icpt->xmmin + ix * icpt->dx - xcent;
        double rsq = xx * xx + yy * yy;
        bool in = outside ? (rsq >= r2) : (rsq <= r2);
        Z.at(1, ix, iy) = in ? zcent + zlo : LARGEPOS;
        Z.at(2, ix, iy) = in ? zcent + zhi : LARGEPOS;
      }
    }
  });
}

int wm_zsrfrv_intercepts(int64_t nn, const double *rsrf, const double *zsrf, double xcent,
                         double ycent, double zcent, wm_intercepts *icpt) {
  WM_ENTRY({
    WM_REQUIRE(icpt && rsrf && zsrf && nn >= 2, "wm_zsrfrv_intercepts: bad arguments");
    int64_t ns = nn - 1;
    WM_REQUIRE(icpt->nxicpt == 2 * ns && icpt->nyicpt == 2 * ns && icpt->nzicpt == 2 * ns,
               "wm_zsrfrv_intercepts: intercept array sizes must be 2*(nn-1)");
    // 1-based segment tables (f3d_conductors.F:7093-7099).
    std::vector<double> R(nn + 1), Z(nn + 1), rmin(ns + 1), rmax(ns + 1), zmin(ns + 1),
        zmax(ns + 1);
    for (int64_t i = 1; i <= nn; i++) {
      R[i] = rsrf[i - 1];
      Z[i] = zsrf[i - 1];
    }
    for (int64_t i = 1; i <= ns; i++) {
      rmin[i] = std::min(R[i], R[i + 1]);
      rmax[i] = std::max(R[i], R[i + 1]);
      zmin[i] = std::min(Z[i], Z[i + 1]);
      zmax[i] = std::max(Z[i], Z[i + 1]);
    }
    // Does a line at coordinate q (z for transverse lines, r for axial
    // lines) cross segment ii, and is the segment flat in that coordinate?
    auto crosses = [&](const std::vector<double> &lo, const std::vector<double> &hi,
                       const std::vector<double> &S, int64_t ii, double q) {
      bool ll = (lo[ii] < q && q < hi[ii]) || q == S[ii];
      if (q == S[ii + 1]) {
        int64_t ip1 = ii == nn - 1 ? 1 : ii + 1;
        if ((lo[ii] < q && lo[ip1] < q) || (hi[ii] > q && hi[ip1] > q)) ll = true;
      }
      return ll;
    };
    auto flat_counts = [&](const std::vector<double> &lo, const std::vector<double> &hi,
                           int64_t ii, double q) {
      int64_t im1 = ii - 1 == 0 ? nn - 1 : ii - 1;
      int64_t ip1 = ii + 1 == nn ? 1 : ii + 1;
      while (lo[ip1] == hi[ip1] && ip1 != ii) {
        ip1 = ip1 + 1;
        if (ip1 == nn) ip1 = 1;
      }
      return (lo[im1] < q && lo[ip1] < q) || (hi[im1] > q && hi[ip1] > q);
    };
    std::vector<int64_t> flats;
    // Transverse lines: L holds lines indexed (a, b) with a the transverse
    // coordinate along the other axis and b the z index.
    auto transverse = [&](const Lines &L, int64_t na, int64_t nb, double amin, double da,
                          double acent, double cent) {
      for (int64_t b = 0; b <= nb; b++) {
        double zz = icpt->zmmin + b * icpt->dz - zcent;
        for (int64_t a = 0; a <= na; a++) {
          double t = amin + a * da - acent;
          double *line = &L.at(1, a, b);
          int64_t n = 0;
          flats.clear();
          for (int64_t ii = 1; ii <= ns; ii++) {
            if (!crosses(zmin, zmax, Z, ii, zz)) continue;
            double rr;
            if (Z[ii + 1] - Z[ii] == 0.) {
              flats.push_back(ii);
              if (!flat_counts(zmin, zmax, ii, zz)) continue;
              rr = R[ii];
            } else {
              double ww = (zz - Z[ii]) / (Z[ii + 1] - Z[ii]);
              rr = R[ii] * (1. - ww) + R[ii + 1] * ww;
            }
            if (t * t <= rr * rr) {
              double h1 = -std::sqrt(std::max(0., rr * rr - t * t));
              double h2 = +std::sqrt(std::max(0., rr * rr - t * t));
              insert_icpt(h1 + cent, n, line);
              insert_icpt(h2 + cent, n, line);
            }
          }
          for (int64_t ii : flats) {
            if (t > std::max(R[ii], R[ii + 1])) continue;
            double rlo = std::min(R[ii], R[ii + 1]), rhi = std::max(R[ii], R[ii + 1]);
            double lo = std::sqrt(std::max(0., rlo * rlo - t * t));
            double hi = std::sqrt(std::max(0., rhi * rhi - t * t));
            flat_fix_transverse(line, n, lo, hi, cent);
          }
          for (int64_t k = n; k < L.n; k++) line[k] = LARGEPOS;
        }
      }
    };
    transverse(xlines(icpt), icpt->ny, icpt->nz, icpt->ymmin, icpt->dy, ycent, xcent);
    transverse(ylines(icpt), icpt->nx, icpt->nz, icpt->xmmin, icpt->dx, xcent, ycent);
    Lines Zl = zlines(icpt);
    for (int64_t iy = 0; iy <= icpt->ny; iy++) {
      double yy = icpt->ymmin + iy * icpt->dy - ycent;
      for (int64_t ix = 0; ix <= icpt->nx; ix++) {
        double xx = icpt->xmmin + ix * icpt->dx - xcent;
        double rr = std::sqrt(xx * xx + yy * yy);
        double *line = &Zl.at(1, ix, iy);
        int64_t n = 0;
        flats.clear();
        for (int64_t ii = 1; ii <= ns; ii++) {
          if (!crosses(rmin, rmax, R, ii, rr)) continue;
          if (R[ii + 1] - R[ii] == 0.) {
            flats.push_back(ii);
            if (flat_counts(rmin, rmax, ii, rr)) insert_icpt(Z[ii] + zcent, n, line);
          } else {
            double ww = (rr - R[ii]) / (R[ii + 1] - R[ii]);
            double zz = Z[ii] * (1. - ww) + Z[ii + 1] * ww;
            insert_icpt(zz + zcent, n, line);
          }
        }
        for (int64_t ii : flats)
          flat_fix_axial(line, n, std::min(Z[ii], Z[ii + 1]) + zcent, std::max(Z[ii], Z[ii + 1]) + zcent);
        for (int64_t k = n; k < Zl.n; k++) line[k] = LARGEPOS;
      }
    }
  });
}

int wm_condlist_create(wm_condlist **out) {
  WM_ENTRY({
    WM_REQUIRE(out != nullptr, "wm_condlist_create: NULL");
    *out = new wm_condlist();
  });
}

int wm_condlist_destroy(wm_condlist *c) {
  WM_ENTRY({ delete c; });
}

int wm_conductordelfromintercepts(const wm_intercepts *icpt, wm_condlist *c, double dfill,
                                  double fuzz) {
  WM_ENTRY({
    WM_REQUIRE(icpt && c, "wm_conductordelfromintercepts: NULL argument");
    if (icpt->nx < 0 || icpt->ny < 0 || icpt->nz < 0) return 0;
    const double xmmin = icpt->xmmin, ymmin = icpt->ymmin, zmmin = icpt->zmmin;
    const double dx = icpt->dx, dy = icpt->dy, dz = icpt->dz;
    const int64_t nx = icpt->nx, ny = icpt->ny, nz = icpt->nz;
    const int64_t ixl = icpt->ix, iyl = icpt->iy, izl = icpt->iz;
    const double xmmax = xmmin + nx * dx, ymmax = ymmin + ny * dy, zmmax = zmmin + nz * dz;
    const double dxi = 1. / dx, dyi = 1. / dy, dzi = 1. / dz;
    Lines X = xlines(icpt), Y = ylines(icpt), Z = zlines(icpt);
    const int64_t level = icpt->mglevel;

    // --- Interior points (f3d_conductors.F:180-267).
    for (int64_t iz = 0; iz <= nz; iz++) {
      double zz = zmmin + iz * dz;
      for (int64_t iy = 0; iy <= ny; iy++) {
        double yy = ymmin + iy * dy;
        for (int64_t ii = 1; ii <= X.n; ii += 2) {
          if (X.at(ii, iy, iz) >= LARGEPOS) break;
          double x1 = X.at(ii, iy, iz), x2 = X.at(ii + 1, iy, iz);
          if (!(x2 >= xmmin && x1 <= xmmax)) continue;
          int64_t ix1 = (int64_t)((std::max(x1 - fuzz, xmmin) - xmmin) * dxi);
          int64_t ix2 = (int64_t)((std::min(x2 + fuzz, xmmax) - xmmin) * dxi);
          if (xmmin + ix1 * dx <= x1 && ix1 < nx) ix1++;
          if (xmmin + (ix2 + 1) * dx <= x2 && ix2 < nx) ix2++;
          for (int64_t ix = ix1; ix <= ix2; ix++) {
            double delx, dely, delz;
            if (dfill < LARGEPOS) {
              double xx = xmmin + ix * dx;
              delx = std::min(xx - x1, x2 - xx);
              dely = LARGEPOS;
              delz = LARGEPOS;
              for (int64_t k = 1; k <= Y.n; k++) {
                dely = std::min(dely, std::fabs(yy - Y.at(k, ix, iz)));
                if (Y.at(k, ix, iz) > yy) break;
              }
              for (int64_t k = 1; k <= Z.n; k++) {
                delz = std::min(delz, std::fabs(zz - Z.at(k, ix, iy)));
                if (Z.at(k, ix, iy) > zz) break;
              }
            } else {
              delx = dely = delz = 0.;
            }
            if (delx < dfill * dx || dely < dfill * dy || delz < dfill * dz) {
              c->int_indx.insert(c->int_indx.end(), {ix + ixl, iy + iyl, iz + izl});
              c->int_volt.push_back(X.vo((ii + 1) / 2, iy, iz));
              c->int_numb.push_back(X.id((ii + 1) / 2, iy, iz));
              c->int_ilevel.push_back(level);
            }
          }
        }
      }
    }

    // Distance from a grid coordinate q to the surfaces on a crossing
    // line, scanning intercepts in order (the y and z scans of the x pass,
    // f3d_conductors.F:337-407).  Updates dels/volt/numb slots m (minus)
    // and p (plus) when a surface is nearer than 1-fuzz.
    auto nearest = [&](const Lines &L, int64_t a, int64_t b, double q, double di, int m, int p,
                       double *dels, double *volt, int64_t *numb) {
      double side = +1.;
      for (int64_t k = 1; k <= L.n; k++) {
        double s = L.at(k, a, b);
        if (s >= LARGEPOS) break;
        double del;
        int id;
        if (side > 0.) {
          if (k == 1 || q > L.at(k - 1, a, b)) del = (s - q) * di;
          else del = LARGEPOS;
          id = p;
        } else {
          if (k == L.n || q < L.at(k + 1, a, b)) del = (q - s) * di;
          else del = LARGEPOS;
          id = m;
        }
        if (0. < del && del < 1. - fuzz && del < dels[id]) {
          dels[id] = del;
          volt[id] = L.vo((k + 1) / 2, a, b);
          numb[id] = L.id((k + 1) / 2, a, b);
        }
        side = -side;
      }
    };
    // Whether a grid point already has a surface within one cell along an
    // earlier-scanned axis (the checks of f3d_conductors.F:478-497).
    auto near_along = [&](const Lines &L, int64_t a, int64_t b, double q, double di) {
      double side = +1., del = 2.;
      for (int64_t k = 1; k <= L.n; k++) {
        double s = L.at(k, a, b);
        if (s >= LARGEPOS) break;
        if (del < 1.) break;
        if (side > 0.) {
          if (s >= q) del = (s - q) * di;
        } else {
          if (s <= q) del = (q - s) * di;
        }
        side = -side;
      }
      return del;
    };
    auto save = [&](int64_t ix, int64_t iy, int64_t iz, const double *dels, const double *volt,
                    const int64_t *numb) {
      int par = (int)((ix + ixl + iy + iyl + iz + izl) % 2);
      SubGrid &s = c->sub[par];
      s.indx.insert(s.indx.end(), {ix + ixl, iy + iyl, iz + izl});
      s.dels.insert(s.dels.end(), dels, dels + 6);
      s.volt.insert(s.volt.end(), volt, volt + 6);
      s.numb.insert(s.numb.end(), numb, numb + 6);
      s.ilevel.push_back(level);
    };
    // The surface crossing of one intercept along its line: the grid index
    // of the outside point next to it, its fractional distance and the
    // dels slot (f3d_conductors.F:293-321).
    auto crossing = [&](const Lines &L, int64_t a, int64_t b, int64_t ii, double side, double mn,
                        double di, double d, int64_t n, int mslot, int pslot, int64_t &i,
                        double &del, int &slot) -> bool {
      double s = L.at(ii, a, b);
      if (!(mn - d < s && s < mn + n * d + d)) return false;
      i = (int64_t)std::floor((s - mn) * di);
      if (side > 0.) {
        if (i < 0) return false;
        double q = mn + i * d;
        if (ii == 1 || q > L.at(ii - 1, a, b)) del = (s - mn) * di - i;
        else del = LARGEPOS;
        slot = pslot;
      } else {
        i = i + 1;
        if (i > n) return false;
        double q = mn + i * d;
        if (ii == L.n || q < L.at(ii + 1, a, b)) del = i - (s - mn) * di;
        else del = LARGEPOS;
        slot = mslot;
      }
      return true;
    };
    (void)xmmax;
    (void)ymmax;
    (void)zmmax;

    // --- Subgrid points from the x intercepts (f3d_conductors.F:274-432).
    for (int64_t iz = 0; iz <= nz; iz++) {
      double zz = zmmin + iz * dz;
      for (int64_t iy = 0; iy <= ny; iy++) {
        double yy = ymmin + iy * dy;
        double side = -1.;
        for (int64_t ii = 1; ii <= X.n; ii++) {
          if (X.at(ii, iy, iz) >= LARGEPOS) break;
          side = -side;
          int64_t ix;
          double del;
          int slot;
          if (!crossing(X, iy, iz, ii, side, xmmin, dxi, dx, nx, 0, 1, ix, del, slot)) continue;
          if (!(0. < del && del < 1. - fuzz)) continue;
          double dels[6], volt[6];
          int64_t numb[6];
          std::fill(dels, dels + 6, LARGEPOS);
          std::fill(volt, volt + 6, 0.);
          std::fill(numb, numb + 6, (int64_t)0);
          dels[slot] = del;
          volt[slot] = X.vo((ii + 1) / 2, iy, iz);
          numb[slot] = X.id((ii + 1) / 2, iy, iz);
          nearest(Y, ix, iz, yy, dyi, 2, 3, dels, volt, numb);
          nearest(Z, ix, iy, zz, dzi, 4, 5, dels, volt, numb);
          save(ix, iy, iz, dels, volt, numb);
        }
      }
    }
    // --- Subgrid points from the y intercepts (f3d_conductors.F:434-584).
    for (int64_t iz = 0; iz <= nz; iz++) {
      double zz = zmmin + iz * dz;
      for (int64_t ix = 0; ix <= nx; ix++) {
        double xx = xmmin + ix * dx;
        double side = -1.;
        for (int64_t ii = 1; ii <= Y.n; ii++) {
          if (Y.at(ii, ix, iz) >= LARGEPOS) break;
          side = -side;
          int64_t iy;
          double del;
          int slot;
          if (!crossing(Y, ix, iz, ii, side, ymmin, dyi, dy, ny, 2, 3, iy, del, slot)) continue;
          double delx = near_along(X, iy, iz, xx, dxi);
          if (!(0. < del && del < 1. - fuzz && delx >= 1. - fuzz)) continue;
          double dels[6], volt[6];
          int64_t numb[6];
          std::fill(dels, dels + 6, LARGEPOS);
          std::fill(volt, volt + 6, 0.);
          std::fill(numb, numb + 6, (int64_t)0);
          dels[slot] = del;
          volt[slot] = Y.vo((ii + 1) / 2, ix, iz);
          numb[slot] = Y.id((ii + 1) / 2, ix, iz);
          nearest(Z, ix, iy, zz, dzi, 4, 5, dels, volt, numb);
          save(ix, iy, iz, dels, volt, numb);
        }
      }
    }
    // --- Subgrid points from the z intercepts (f3d_conductors.F:586-719).
    for (int64_t iy = 0; iy <= ny; iy++) {
      double yy = ymmin + iy * dy;
      for (int64_t ix = 0; ix <= nx; ix++) {
        double xx = xmmin + ix * dx;
        double side = -1.;
        for (int64_t ii = 1; ii <= Z.n; ii++) {
          if (Z.at(ii, ix, iy) >= LARGEPOS) break;
          side = -side;
          int64_t iz;
          double del;
          int slot;
          if (!crossing(Z, ix, iy, ii, side, zmmin, dzi, dz, nz, 4, 5, iz, del, slot)) continue;
          double delx = near_along(X, iy, iz, xx, dxi);
          double dely = near_along(Y, ix, iz, yy, dyi);
          if (!(0. < del && del < 1. - fuzz && delx >= 1. - fuzz && dely >= 1. - fuzz)) continue;
          double dels[6], volt[6];
          int64_t numb[6];
          std::fill(dels, dels + 6, LARGEPOS);
          std::fill(volt, volt + 6, 0.);
          std::fill(numb, numb + 6, (int64_t)0);
          dels[slot] = del;
          volt[slot] = Z.vo((ii + 1) / 2, ix, iy);
          numb[slot] = Z.id((ii + 1) / 2, ix, iy);
          save(ix, iy, iz, dels, volt, numb);
        }
      }
    }
  });
}

int wm_checkconductors(wm_condlist *c, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                       double dz) {
  WM_ENTRY({
    WM_REQUIRE(c != nullptr, "wm_checkconductors: NULL list");
    mglevels(c, nx, ny, nz, dx, dy, dz);
    for (int64_t l = 0; l < c->levels; l++)
      clean_level(c, c->levelnx[l], c->levelny[l], c->levelnz[l], l);
    std::vector<int64_t> order = level_order(c->int_ilevel, c->levels, c->int_istart);
    gather_rows(c->int_indx, order, 3);
    gather_rows(c->int_volt, order, 1);
    gather_rows(c->int_numb, order, 1);
    gather_rows(c->int_ilevel, order, 1);
    for (int p = 0; p < 2; p++) {
      SubGrid &s = c->sub[p];
      order = level_order(s.ilevel, c->levels, s.istart);
      gather_rows(s.indx, order, 3);
      gather_rows(s.dels, order, 6);
      gather_rows(s.volt, order, 6);
      gather_rows(s.numb, order, 6);
      gather_rows(s.ilevel, order, 1);
    }
  });
}

int wm_condlist_view(const wm_condlist *c, wm_conductor_lists *v) {
  WM_ENTRY({
    WM_REQUIRE(c && v, "wm_condlist_view: NULL argument");
    WM_REQUIRE(c->levels > 0, "wm_condlist_view: run wm_checkconductors first");
    v->levels = c->levels;
    v->int_n = (int64_t)c->int_volt.size();
    v->int_indx = c->int_indx.data();
    v->int_volt = c->int_volt.data();
    v->int_istart = c->int_istart.data();
    for (int p = 0; p < 2; p++) {
      v->sub_n[p] = c->sub[p].n();
      v->sub_indx[p] = c->sub[p].indx.data();
      v->sub_dels[p] = c->sub[p].dels.data();
      v->sub_volt[p] = c->sub[p].volt.data();
      v->sub_istart[p] = c->sub[p].istart.data();
    }
  });
}

}  // extern "C"
