// Harness around the reference's own neighbour-search, kernel, convex-hull
// and k-d tree sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from
// the reference tree as it lies on disk (src/kernel.cpp, src/particle.cpp,
// src/kd_tree.cpp, src/grid.h, src/convex_hull.h,
// src/benchmarks/neighbor_search.h) into oracle/_ref/, and used
// only to pin the oracle and to write the golden vectors under tests/golden.
// The reference's physics headers (derivatives.h, correctors.h, material.h,
// cont_mech.h, precomp_shape_functions.h, body.h, rk4.h) include glm, which
// is absent from this image, so they are not built.
#include <stdio.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "convex_hull.h"
#include "grid.h"
#include "kd_tree.h"
#include "kernel.h"
#include "particle.h"
#include "benchmarks/neighbor_search.h"

extern "C" {

void ref_cubic_spline(double xi, double yi, double xj, double yj, double h, double out[3]) {
  kernel_result w = cubic_spline(xi, yi, xj, yj, h);
  out[0] = w.w;
  out[1] = w.w_x;
  out[2] = w.w_y;
}

// One UL stage of neighbour search exactly as src/body.h:134-138 runs it:
// update_geometry, assign_hashes, std::sort by hash, construct_verlet_lists.
// Lists are returned in idx space, rows by idx, in emission order.
// geom_out: bbmin_x, bbmin_y, cell side, nx, ny.
int ref_verlet_pairs(uint32_t n, const double* x, const double* y, const double* h, uint32_t* row_ptr,
                     uint32_t* col, uint64_t cap, double* geom_out) {
  std::vector<particle*> ps(n);
  for (uint32_t i = 0; i < n; i++) {
    ps[i] = new particle(i);
    ps[i]->x = x[i];
    ps[i]->y = y[i];
    ps[i]->h = h[i];
  }
  particle** P = ps.data();
  grid<particle> g;
  g.update_geometry(P, n, 2.);
  g.assign_hashes(P, n);
  std::sort(P, P + n, [](const particle* a, const particle* b) { return a->hash < b->hash; });
  g.construct_verlet_lists(P, n, 2.);
  std::vector<uint32_t> cnt(n);
  for (uint32_t i = 0; i < n; i++) cnt[P[i]->idx] = P[i]->num_nbh;
  row_ptr[0] = 0;
  for (uint32_t k = 0; k < n; k++) row_ptr[k + 1] = row_ptr[k] + cnt[k];
  int rc = 0;
  if (row_ptr[n] > cap) rc = -9;
  for (uint32_t i = 0; i < n && rc == 0; i++) {
    uint32_t o = row_ptr[P[i]->idx];
    for (uint32_t q = 0; q < P[i]->num_nbh; q++) col[o + q] = P[P[i]->nbh[q]]->idx;
  }
  if (geom_out) {
    // recover the cell geometry from the hashes the reference assigned
    double bminx = 1e300, bminy = 1e300, hmax = -1e300;
    for (uint32_t i = 0; i < n; i++) {
      bminx = fmin(bminx, x[i]);
      bminy = fmin(bminy, y[i]);
      hmax = fmax(hmax, h[i]);
    }
    geom_out[0] = bminx - 1e-6;
    geom_out[1] = bminy - 1e-6;
    geom_out[2] = hmax * 2.;
  }
  for (auto* p : ps) delete p;
  return rc;
}

// src/benchmarks/neighbor_search.h:116-147 (TL initialisation).
int ref_find_neighbors(uint32_t n, const double* x, const double* y, double h, uint32_t* row_ptr,
                       uint32_t* col, uint64_t cap) {
  std::vector<particle*> ps(n);
  for (uint32_t i = 0; i < n; i++) {
    ps[i] = new particle(i);
    ps[i]->x = x[i];
    ps[i]->y = y[i];
    ps[i]->h = h;
  }
  find_neighbors(ps.data(), n, h, distance_euclidian);
  row_ptr[0] = 0;
  int rc = 0;
  for (uint32_t i = 0; i < n; i++) {
    row_ptr[i + 1] = row_ptr[i] + ps[i]->num_nbh;
    if (row_ptr[i + 1] > cap) {
      rc = -9;
      break;
    }
    for (uint32_t q = 0; q < ps[i]->num_nbh; q++) col[row_ptr[i] + q] = ps[i]->nbh[q];
  }
  for (auto* p : ps) delete p;
  return rc;
}

// src/convex_hull.h:62-95 on particles at (x, y); returns the hull's
// particle indexes and their count (idx_out holds up to n entries).
uint32_t ref_convex_hull(const double* x, const double* y, uint32_t n, uint32_t* idx_out) {
  std::vector<particle*> ps(n);
  for (uint32_t i = 0; i < n; i++) {
    ps[i] = new particle(i);
    ps[i]->x = x[i];
    ps[i]->y = y[i];
  }
  std::vector<unsigned int> h = convex_hull(ps.data(), n);
  for (size_t k = 0; k < h.size() && k < n; k++) idx_out[k] = h[k];
  for (auto* p : ps) delete p;
  return (uint32_t)h.size();
}

// kd_tree (src/kd_tree.cpp:20-116) over points pts (x, y, nx, ny per point):
// the closest point of each query, written as (x, y, nx, ny).
void ref_kd_nearest(const double* pts, uint32_t nb, const double* qx, const double* qy, uint32_t nq, double* out) {
  std::vector<kd_point> p(nb);
  for (uint32_t k = 0; k < nb; k++) p[k] = kd_point{pts[4 * k], pts[4 * k + 1], pts[4 * k + 2], pts[4 * k + 3]};
  kd_tree tree(p.data(), nb);
  for (uint32_t i = 0; i < nq; i++) {
    kd_point c = tree.find_closest_point(qx[i], qy[i]);
    out[4 * i] = c.x;
    out[4 * i + 1] = c.y;
    out[4 * i + 2] = c.nx;
    out[4 * i + 3] = c.ny;
  }
}

}  // extern "C"
